// path_cache.hip -- the path cache: every LOR's Siddon path walked once and
// kept in HBM, so that projector passes stream the system matrix instead of
// walking it again.
//
// Two copies of the matrix are kept:
//   - the path matrix (RowMatrix), LOR-major: the LORs that cross the
//     volume ("rows") in (view, segment, axial, tangential) order, each
//     with its elements (voxel index, length) in path order. Forward
//     projections and OSEM line integrals are gathers over it, T lanes per
//     row.
//   - the transpose of one subset split (fir_cache), voxel-major: for
//     every subset and voxel the (row, length) elements of the subset's
//     LORs through that voxel, in row order. Backprojection is a gather
//     over it, T lanes per voxel: no atomics, and the result does not
//     depend on scheduling.
// The reference's LORCache (LORCache.h:5-47, LORCache.cc:10-168) keeps the
// crystal pairs of every (subset, segment) to skip their decoding; this
// cache keeps whole paths, which the 288 GB of HBM of one GPU hold for a
// clinical scanner (3.9e9 elements, 23 GB per copy at 6 B per element).
//
// Element lengths are stored as
//   FIR_LEN_SNORM16: round(32767 * len / m), m = the LOR's largest |length|
//                    (the reference walk with int-truncated steps makes some
//                    lengths negative);
//   FIR_LEN_FP16:    binary16 of len / m;
//   FIR_LEN_FP32:    binary32 of len;
// with a per-row FP32 factor (m / 32767, m, 1) that restores the length.
#include <hip/hip_fp16.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>

#include "internal.h"

using fir::Grid;
using fir::kEpsF;
using fir::Lor;
using fir::LorGeom;
using fir::Walk;
using namespace firi;

namespace
{

constexpr int kMaxCacheBatch = 8;

// Elements per transpose chunk (bounds the sort's scratch: 24 B each).

template <int F>
struct Len;

template <>
struct Len<FIR_LEN_SNORM16>
{
  using T = short;
  static __device__ __forceinline__ float dec(T v) { return (float)v; }
  static __device__ __forceinline__ T enc(double len, float m)
  {
    if (!(m > 0.0f)) return 0;
    double q = rint(len / (double)m * 32767.0);
    q = q < -32767.0 ? -32767.0 : (q > 32767.0 ? 32767.0 : q);
    return (T)q;
  }
  static __device__ __forceinline__ unsigned bits(T v)
  {
    return (unsigned)(unsigned short)v;
  }
  static __device__ __forceinline__ T from_bits(unsigned b)
  {
    return (T)(unsigned short)b;
  }
};

template <>
struct Len<FIR_LEN_FP16>
{
  using T = __half;
  static __device__ __forceinline__ float dec(T v) { return __half2float(v); }
  static __device__ __forceinline__ T enc(double len, float m)
  {
    if (!(m > 0.0f)) return __float2half_rn(0.0f);
    return __float2half_rn((float)(len / (double)m));
  }
  static __device__ __forceinline__ unsigned bits(T v)
  {
    return (unsigned)__half_as_ushort(v);
  }
  static __device__ __forceinline__ T from_bits(unsigned b)
  {
    return __ushort_as_half((unsigned short)b);
  }
};

template <>
struct Len<FIR_LEN_FP32>
{
  using T = float;
  static __device__ __forceinline__ float dec(T v) { return v; }
  static __device__ __forceinline__ T enc(double len, float) { return (float)len; }
  static __device__ __forceinline__ unsigned bits(T v) { return __float_as_uint(v); }
  static __device__ __forceinline__ T from_bits(unsigned b) { return __uint_as_float(b); }
};

size_t len_bytes(int fmt) { return fmt == FIR_LEN_FP32 ? 4 : 2; }

// Image layouts of the gathers. The projector kernels gather from two
// copies of each image: x fastest for LORs closer to the x axis, y fastest
// for the others, so that the consecutive voxels of a path share cache
// lines. In both, zq consecutive z-slices are interleaved innermost:
// voxel (x, y, z) of the x-fastest copy sits at
// (((z / zq) ny + y) nx + x) zq + z % zq (x and y swapped in the other).
// The rows a wave gathers for together are LORs through one line of the
// xy-plane at neighbouring centre slices, a few slices apart, whose voxels
// then share cache lines too (zq is fixed per path matrix,
// fir_set_cache_zslices).
struct XY
{
  int nx, ny;
  int zq = 1;
};

// Slices of the copies: nz rounded up to a multiple of zq.
long long layout_voxels(const XY& g, int nz)
{
  return (long long)g.nx * g.ny * ((nz + g.zq - 1) / g.zq * g.zq);
}

__device__ __forceinline__ int to_layout(const XY& g, int idx, bool yf)
{
  const int x = idx % g.nx;
  const int r = idx / g.nx;
  const int y = r % g.ny;
  const int z = r / g.ny;
  const int a = yf ? y : x, b = yf ? x : y;
  const int na = yf ? g.ny : g.nx, nb = yf ? g.nx : g.ny;
  return (((z / g.zq) * nb + b) * na + a) * g.zq + z % g.zq;
}

__device__ __forceinline__ int from_layout(const XY& g, int li, bool yf)
{
  const int na = yf ? g.ny : g.nx, nb = yf ? g.nx : g.ny;
  const int zl = li % g.zq;
  const int r = li / g.zq;
  const int a = r % na;
  const int r2 = r / na;
  const int b = r2 % nb;
  const int z = (r2 / nb) * g.zq + zl;
  const int x = yf ? b : a, y = yf ? a : b;
  return (z * g.ny + y) * g.nx + x;
}

__device__ __forceinline__ bool y_dominant(const Lor& L)
{
  const double dx = L.e2[0] - L.e1[0], dy = L.e2[1] - L.e1[1];
  return (dy < 0 ? -dy : dy) > (dx < 0 ? -dx : dx);
}

// ---------------------------------------------------------------------------
// Build kernels
// ---------------------------------------------------------------------------

// LOR t of the cache's row order: view by view (t / per_view); within a
// view the decode of `lg` for a one-view subset -- (pair, tangential) or,
// with lg.tang_major, (tangential, pair) -- with the pairs in lg's order.
__device__ __forceinline__ void cache_lor(const LorGeom& lg, long long t,
                                          long long per_view, Lor& L)
{
  const int view = (int)(t / per_view);
  fir::lor_from_index(lg, t - (long long)view * per_view, lg.n_views, view, L);
}

// Pass 1 over all LORs in the cache's row order: element count and largest
// |length| of each path.
__global__ __launch_bounds__(kBlock) void k_pc_stats(Grid g, LorGeom lg,
                                                     long long n,
                                                     long long per_view,
                                                     int* __restrict__ count,
                                                     float* __restrict__ maxlen)
{
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  Lor L;
  cache_lor(lg, t, per_view, L);
  Walk w;
  int c = 0;
  double m = 0.0;
  if (fir::walk_setup(g, L.e1, L.e2, w))
    fir::walk_run(g, w, [&](int, double l) {
      ++c;
      const double a = l < 0 ? -l : l;
      m = a > m ? a : m;
    });
  count[t] = c;
  maxlen[t] = (float)m;
}

struct HitFlag
{
  __host__ __device__ int operator()(int c) const { return c > 0 ? 1 : 0; }
};

struct ToI64
{
  __host__ __device__ long long operator()(int c) const { return (long long)c; }
  __host__ __device__ long long operator()(unsigned c) const
  {
    return (long long)c;
  }
};

// Element slots of a path: its count rounded up to a multiple of 4, so that
// every path (and every voxel's transposed list) starts 16-byte aligned and
// a lane loads four consecutive elements at once; the padding elements
// have index 0 and length 0.
struct Pad4
{
  __host__ __device__ long long operator()(int c) const
  {
    return (long long)((c + 3) & ~3);
  }
  __host__ __device__ long long operator()(unsigned c) const
  {
    return (long long)((c + 3u) & ~3u);
  }
};

// Slot of element p of a list of n slots (n a multiple of 4) that T lanes
// read quad by quad (lane j takes quads j, j + T, ...): within each round of
// T quads, element k of lane j's quad holds list element k * r + j of the
// round (r = quads in the round). Each load instruction of the T lanes
// then reads consecutive list elements -- consecutive voxels of a path,
// consecutive rows of a voxel -- so their gathers share cache lines.
__host__ __device__ __forceinline__ long long quad_slot(long long p,
                                                        long long n, int T)
{
  const long long q = n >> 2;
  const long long m = p / (4LL * T);
  const long long pl = p - m * 4LL * T;
  const long long left = q - m * T;
  const long long r = left < T ? left : T;
  const long long k = pl / r;
  const long long j = pl - k * r;
  return 4 * (m * T + j) + k;
}

// Row tables: bin, length factor and first element of every crossing LOR.
__global__ __launch_bounds__(kBlock) void k_pc_rows(
  LorGeom lg, long long n, long long per_view, const int* __restrict__ count,
  const float* __restrict__ maxlen, const int* __restrict__ rowid,
  const long long* __restrict__ eoff, int fmt, int* __restrict__ bin,
  float* __restrict__ factor, long long* __restrict__ off,
  unsigned char* __restrict__ yrow)
{
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n || count[t] == 0) return;
  Lor L;
  cache_lor(lg, t, per_view, L);
  const int r = rowid[t];
  const float m = maxlen[t];
  bin[r] = (int)L.bin;
  yrow[r] = y_dominant(L) ? 1 : 0;
  off[r] = eoff[t];
  factor[r] = fmt == FIR_LEN_SNORM16 ? m / 32767.0f
            : fmt == FIR_LEN_FP16    ? m
                                     : 1.0f;
}

// First row and first element of every view, and the totals.
__global__ void k_pc_views(const int* __restrict__ count,
                           const int* __restrict__ rowid,
                           const long long* __restrict__ eoff, long long n,
                           long long per_view, int n_views,
                           int* __restrict__ view_row,
                           long long* __restrict__ view_elem,
                           long long* __restrict__ row_off_end)
{
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > n_views) return;
  if (v < n_views)
  {
    view_row[v] = rowid[v * per_view];
    view_elem[v] = eoff[v * per_view];
  }
  else
  {
    const int rows = rowid[n - 1] + (count[n - 1] > 0 ? 1 : 0);
    const long long elems = eoff[n - 1] + ((count[n - 1] + 3) & ~3);
    view_row[v] = rows;
    view_elem[v] = elems;
    row_off_end[rows] = elems; // off[R] = E
  }
}

// Element slots, path elements and crossing LORs of each view (one block
// per view): the device memory each subset's share of a cache takes.
__global__ __launch_bounds__(kBlock) void k_pc_view_totals(
  const int* __restrict__ count, long long per_view,
  unsigned long long* __restrict__ totals) // [n_views][3]
{
  const int* c = count + (long long)blockIdx.x * per_view;
  unsigned long long t[3] = {0, 0, 0};
  for (long long i = threadIdx.x; i < per_view; i += kBlock)
  {
    const int k = c[i];
    t[0] += (unsigned)((k + 3) & ~3);
    t[1] += (unsigned)k;
    t[2] += k > 0 ? 1u : 0u;
  }
  __shared__ unsigned long long sh[3][kBlock];
  for (int j = 0; j < 3; ++j) sh[j][threadIdx.x] = t[j];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1)
  {
    if ((int)threadIdx.x < s)
      for (int j = 0; j < 3; ++j) sh[j][threadIdx.x] += sh[j][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3)
    totals[3LL * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// A partial cache holds subsets 0 .. k-1 of S: the LORs of the other
// subsets' views (view % S >= k, LORCache.cc:49-51) get no row.
__global__ void k_pc_keep_subsets(int* __restrict__ count, long long n,
                                  long long per_view, int S, int k)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && (int)(t / per_view) % S >= k) count[t] = 0;
}

// sino[b] = 0 for the bins of the views of subsets 0 .. k-1 of S; the
// sinogram is [segment][view][axial][tangential] (ProjData.h:16-21).
__global__ void k_pc_zero_subset_bins(float* __restrict__ sino, long long n_bins,
                                      const long long* __restrict__ seg_base,
                                      const int* __restrict__ n_axial,
                                      int n_seg, int n_tang, int S, int k)
{
  for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       b < n_bins; b += (long long)gridDim.x * blockDim.x)
  {
    int lo = 0, hi = n_seg - 1; // last segment starting at or before b
    while (lo < hi)
    {
      const int mid = (lo + hi + 1) / 2;
      if (seg_base[mid] <= b) lo = mid;
      else hi = mid - 1;
    }
    const long long view = (b - seg_base[lo]) / n_tang / n_axial[lo];
    if ((int)(view % S) < k) sino[b] = 0.0f;
  }
}

// Four lengths of a quad as one store.
template <int F>
__device__ __forceinline__ void store_len4(typename Len<F>::T* p,
                                           const typename Len<F>::T* v)
{
  if constexpr (sizeof(typename Len<F>::T) == 2)
    *reinterpret_cast<uint2*>(p) =
      make_uint2(Len<F>::bits(v[0]) | (Len<F>::bits(v[1]) << 16),
                 Len<F>::bits(v[2]) | (Len<F>::bits(v[3]) << 16));
  else
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}

// Pass 2: the paths themselves. With stage set (lanes <= 8), a thread
// collects each round of its path's quad layout (4 * r elements, r the
// round's quads) in its own LDS slots and writes the round out as whole
// quads, padding included: a round's slots are contiguous, so the stores
// fill whole cache lines at once instead of leaving lines of many rows
// half written in L2 (the elements arrive in path order, which the
// interleave scatters over the round). Dynamic LDS: kBlock * 4 * lanes
// indices and as many lengths.
template <int F>
__global__ __launch_bounds__(kBlock) void k_pc_fill(
  Grid g, LorGeom lg, long long n, long long per_view,
  const int* __restrict__ count,
  const float* __restrict__ maxlen, const long long* __restrict__ eoff,
  XY xy, int lanes, int stage, int* __restrict__ idx,
  typename Len<F>::T* __restrict__ len)
{
  using T = typename Len<F>::T;
  extern __shared__ int fill_lds[];
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n || count[t] == 0) return;
  Lor L;
  cache_lor(lg, t, per_view, L);
  Walk w;
  if (!fir::walk_setup(g, L.e1, L.e2, w)) return;
  const float m = maxlen[t];
  const bool yf = y_dominant(L);
  const long long base = eoff[t];
  const int cnt = count[t];
  const long long slots = (cnt + 3) & ~3;
  int p = 0;
  if (!stage)
  {
    fir::walk_run(g, w, [&](int i, double l) {
      if (p < cnt)
      {
        const long long e = base + quad_slot(p, slots, lanes);
        idx[e] = to_layout(xy, i, yf);
        len[e] = Len<F>::enc(l, m);
      }
      ++p;
    });
    return;
  }
  int* si = fill_lds + threadIdx.x;
  T* sl = reinterpret_cast<T*>(fill_lds + 4 * lanes * kBlock) + threadIdx.x;
  const int q = (int)(slots >> 2);
  int rnd = 0, pl = 0;
  int r = q < lanes ? q : lanes;
  fir::walk_run(g, w, [&](int i, double l) {
    if (p < cnt)
    {
      // element pl of the round: element k * r + j of the round sits in
      // slot 4 j + k (quad_slot)
      const int k = pl / r, j = pl - k * r;
      si[(4 * j + k) * kBlock] = to_layout(xy, i, yf);
      sl[(4 * j + k) * kBlock] = Len<F>::enc(l, m);
      ++pl;
      if (pl == 4 * r || p + 1 == cnt)
      {
        const long long b0 = base + 4LL * rnd * lanes;
        for (int jj = 0; jj < r; ++jj)
        {
          int v[4];
          T lv[4];
#pragma unroll
          for (int kk = 0; kk < 4; ++kk)
          {
            const bool real = kk * r + jj < pl;
            v[kk] = real ? si[(4 * jj + kk) * kBlock] : 0;
            lv[kk] = real ? sl[(4 * jj + kk) * kBlock] : (T)0;
          }
          *reinterpret_cast<int4*>(idx + b0 + 4 * jj) =
            make_int4(v[0], v[1], v[2], v[3]);
          store_len4<F>(len + b0 + 4 * jj, lv);
        }
        ++rnd;
        pl = 0;
        const int left = q - rnd * lanes;
        r = left < lanes ? left : lanes;
      }
    }
    ++p;
  });
}

// Rows of subset s: view k * S + s for k = blockIdx.y. With xcd_map, the
// x-blocks of a view (a multiple of 8) are dealt so that the XCD that runs
// block x (x mod 8, as the dispatcher deals workgroups round-robin) gets the
// contiguous eighth (x mod 8) of the view's rows: with rows ordered by
// centre slice, each XCD's L2 then serves one z-slab of the image.
struct SubsetRows
{
  const int* view_row;
  int n_subsets;
  int subset;
  int xcd_map;
};

__device__ __forceinline__ unsigned logical_block(unsigned b, unsigned n,
                                                  int xcd_map)
{
  return xcd_map ? (b & 7u) * (n >> 3) + (b >> 3) : b;
}

__device__ __forceinline__ bool subset_row(const SubsetRows& sr, int lanes,
                                           int* r, int* r_first, int* k)
{
  const int view = (int)blockIdx.y * sr.n_subsets + sr.subset;
  const int r0 = sr.view_row[view];
  const int r1 = sr.view_row[view + 1];
  const unsigned bx = logical_block(blockIdx.x, gridDim.x, sr.xcd_map);
  *k = (int)blockIdx.y;
  *r_first = r0;
  *r = r0 + (int)(((long long)bx * blockDim.x + threadIdx.x) / lanes);
  return *r < r1;
}

// Whether element e has a zero length (the padding elements, and elements
// that contribute nothing): the transpose skips them. lb = bytes per length.
__device__ __forceinline__ bool zero_len(const void* len, int lb, long long e)
{
  return lb == 2 ? ((const unsigned short*)len)[e] == 0
                 : ((const unsigned*)len)[e] == 0;
}

// The transpose passes read each row with kTrLanes lanes: lane j takes
// slots j, j + kTrLanes, ... of the row, so that the group's loads are
// contiguous. Rows are whole per group (groups never straddle a row).
constexpr int kTrLanes = 16;

// Slot e's voxel when it holds a contributing element (length != 0), else -1.
__device__ __forceinline__ int tr_voxel(const int* __restrict__ idx,
                                        const void* __restrict__ len, int lb,
                                        const XY& xy, bool yf, long long e,
                                        long long end)
{
  if (e >= end || zero_len(len, lb, e)) return -1;
  return from_layout(xy, idx[e], yf);
}

// Voxel histogram of one subset's elements (indices of y-fastest rows
// converted back).
__global__ __launch_bounds__(kBlock) void k_tr_hist(
  SubsetRows sr, const long long* __restrict__ off, const int* __restrict__ idx,
  const void* __restrict__ len, int lb, const unsigned char* __restrict__ yrow,
  XY xy, unsigned* __restrict__ cnt)
{
  int r, r0, k;
  if (!subset_row(sr, kTrLanes, &r, &r0, &k)) return;
  const int j = threadIdx.x % kTrLanes;
  const bool yf = yrow[r] != 0;
  const long long end = off[r + 1];
  for (long long e = off[r] + j; e - j < end; e += kTrLanes)
  {
    const int v = tr_voxel(idx, len, lb, xy, yf, e, end);
    if (v >= 0) atomicAdd(cnt + v, 1u);
  }
}

// Elements of each row of the subset inside the voxel range [v0, v1);
// local row = row_prefix[k] + (r - first row of view k).
__global__ __launch_bounds__(kBlock) void k_tr_count(
  SubsetRows sr, const int* __restrict__ row_prefix,
  const long long* __restrict__ off, const int* __restrict__ idx,
  const void* __restrict__ len, int lb, const unsigned char* __restrict__ yrow,
  XY xy, int v0, int v1, int* __restrict__ rc)
{
  int r, r0, k;
  if (!subset_row(sr, kTrLanes, &r, &r0, &k)) return;
  const int j = threadIdx.x % kTrLanes;
  int c = 0;
  const unsigned span = (unsigned)(v1 - v0);
  const bool yf = yrow[r] != 0;
  const long long end = off[r + 1];
  for (long long e = off[r] + j; e - j < end; e += kTrLanes)
  {
    const int v = tr_voxel(idx, len, lb, xy, yf, e, end);
    c += v >= 0 && (unsigned)(v - v0) < span ? 1 : 0;
  }
#pragma unroll
  for (int d = kTrLanes / 2; d > 0; d >>= 1) c += __shfl_xor(c, d, kTrLanes);
  if (j == 0) rc[row_prefix[k] + (r - r0)] = c;
}

// (voxel - v0, row | length bits) of each element of the subset inside
// [v0, v1), at the row's positions pos[...]: the group's kept elements in
// slot order. Within a row every voxel appears once, so the stable sort
// by voxel that follows orders each voxel's elements by row whatever the
// order within a row.
template <int F>
__global__ __launch_bounds__(kBlock) void k_tr_expand(
  SubsetRows sr, const int* __restrict__ row_prefix,
  const long long* __restrict__ off, const int* __restrict__ idx,
  const unsigned char* __restrict__ yrow, XY xy,
  const typename Len<F>::T* __restrict__ len, int v0, int v1,
  const int* __restrict__ pos, long long n_out, unsigned* __restrict__ keys,
  unsigned long long* __restrict__ vals)
{
  int r, r0, k;
  if (!subset_row(sr, kTrLanes, &r, &r0, &k)) return;
  const int lane = threadIdx.x % warpSize;
  const int j = lane % kTrLanes;
  const int g0 = lane - j; // first lane of the group
  long long p = pos[row_prefix[k] + (r - r0)];
  const unsigned span = (unsigned)(v1 - v0);
  const bool yf = yrow[r] != 0;
  const long long end = off[r + 1];
  for (long long e = off[r] + j; e - j < end; e += kTrLanes)
  {
    const int v = tr_voxel(idx, len, (int)sizeof(typename Len<F>::T), xy, yf,
                           e, end);
    const unsigned local = (unsigned)(v - v0);
    const bool keep = v >= 0 && local < span;
    const unsigned long long grp =
      (__ballot(keep) >> g0) & ((1ULL << kTrLanes) - 1);
    const long long q = p + __popcll(grp & ((1ULL << j) - 1));
    if (keep && q < n_out)
    {
      keys[q] = local;
      vals[q] = ((unsigned long long)(unsigned)r << 32) |
        (unsigned long long)Len<F>::bits(len[e]);
    }
    p += __popcll(grp);
  }
}

// cnt[v] = elements of voxel v in the sorted keys (cnt[nvox] = 0): two
// binary searches per voxel.
__global__ void k_tr_counts_sorted(const unsigned* __restrict__ keys,
                                   long long n, int nvox,
                                   unsigned* __restrict__ cnt)
{
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > nvox) return;
  long long b[2];
  for (int j = 0; j < 2; ++j)
  {
    const unsigned key = (unsigned)(v + j);
    long long lo = 0, hi = n; // first position with keys[i] >= key
    while (lo < hi)
    {
      const long long mid = (lo + hi) / 2;
      if (keys[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    b[j] = lo;
  }
  cnt[v] = v == nvox ? 0u : (unsigned)(b[1] - b[0]);
}

// Sorted elements of voxels [v0, v0 + span) to their padded slots: the
// element of rank i among voxel v's goes to voff_p[v] + quad_slot(i), where
// the voxel's first rank in the sorted chunk is voff_u[v] - voff_u[v0].
template <int F>
__global__ void k_tr_split(const unsigned* __restrict__ keys,
                           const unsigned long long* __restrict__ vals,
                           long long n, int v0,
                           const long long* __restrict__ voff_p,
                           const long long* __restrict__ voff_u, int lanes,
                           int* __restrict__ trow,
                           typename Len<F>::T* __restrict__ tlen)
{
  const long long u0 = voff_u[v0];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
  {
    const int v = v0 + (int)keys[i];
    const long long dst =
      voff_p[v] + quad_slot(i - (voff_u[v] - u0), voff_p[v + 1] - voff_p[v],
                            lanes);
    const unsigned long long val = vals[i];
    trow[dst] = (int)(unsigned)(val >> 32);
    tlen[dst] = Len<F>::from_bits((unsigned)(val & 0xffffffffULL));
  }
}

// ---------------------------------------------------------------------------
// Projector kernels
// ---------------------------------------------------------------------------

template <int NB>
struct Vec;
template <>
struct Vec<1>
{
  float v[1];
  static __device__ __forceinline__ Vec load(const float* p, long long i)
  {
    Vec r;
    r.v[0] = p[i];
    return r;
  }
};
template <>
struct Vec<2>
{
  float v[2];
  static __device__ __forceinline__ Vec load(const float* p, long long i)
  {
    const float2 a = reinterpret_cast<const float2*>(p)[i];
    Vec r;
    r.v[0] = a.x;
    r.v[1] = a.y;
    return r;
  }
};
template <>
struct Vec<4>
{
  float v[4];
  static __device__ __forceinline__ Vec load(const float* p, long long i)
  {
    const float4 a = reinterpret_cast<const float4*>(p)[i];
    Vec r;
    r.v[0] = a.x;
    r.v[1] = a.y;
    r.v[2] = a.z;
    r.v[3] = a.w;
    return r;
  }
};
template <>
struct Vec<8>
{
  float v[8];
  static __device__ __forceinline__ Vec load(const float* p, long long i)
  {
    const float4 a = reinterpret_cast<const float4*>(p)[2 * i];
    const float4 b = reinterpret_cast<const float4*>(p)[2 * i + 1];
    Vec r;
    r.v[0] = a.x;
    r.v[1] = a.y;
    r.v[2] = a.z;
    r.v[3] = a.w;
    r.v[4] = b.x;
    r.v[5] = b.y;
    r.v[6] = b.z;
    r.v[7] = b.w;
    return r;
  }
};

// Four consecutive lengths (8 or 16 bytes, aligned): raw load, then
// decode to floats.
template <int F>
struct Quad;

template <>
struct Quad<FIR_LEN_SNORM16>
{
  using Raw = uint2;
  static __device__ __forceinline__ Raw load(const void* len, long long e)
  {
    return *reinterpret_cast<const uint2*>((const short*)len + e);
  }
  static __device__ __forceinline__ void dec(Raw q, float* l)
  {
    l[0] = (float)(short)(q.x & 0xffffu);
    l[1] = (float)(short)(q.x >> 16);
    l[2] = (float)(short)(q.y & 0xffffu);
    l[3] = (float)(short)(q.y >> 16);
  }
};

template <>
struct Quad<FIR_LEN_FP16>
{
  using Raw = uint2;
  static __device__ __forceinline__ Raw load(const void* len, long long e)
  {
    return *reinterpret_cast<const uint2*>((const unsigned short*)len + e);
  }
  static __device__ __forceinline__ void dec(Raw q, float* l)
  {
    l[0] = __half2float(__ushort_as_half((unsigned short)(q.x & 0xffffu)));
    l[1] = __half2float(__ushort_as_half((unsigned short)(q.x >> 16)));
    l[2] = __half2float(__ushort_as_half((unsigned short)(q.y & 0xffffu)));
    l[3] = __half2float(__ushort_as_half((unsigned short)(q.y >> 16)));
  }
};

template <>
struct Quad<FIR_LEN_FP32>
{
  using Raw = float4;
  static __device__ __forceinline__ Raw load(const void* len, long long e)
  {
    return *reinterpret_cast<const float4*>((const float*)len + e);
  }
  static __device__ __forceinline__ void dec(Raw q, float* l)
  {
    l[0] = q.x;
    l[1] = q.y;
    l[2] = q.z;
    l[3] = q.w;
  }
};

// acc[b] += sum over the elements of [e, e1) that this lane owns of
// dec(len[i]) * X[ind[i] * NB + b]. e and e1 are multiples of 4 (every path
// and voxel list is padded to whole quads); the lane owns the quads at
// e, e + 4 lanes, ... and loads each with one 16-byte index load and one
// 8-byte length load. With one or two images two quads are in flight per
// trip, with more images one quad (the gathers of 4+ images per element
// would cost registers and occupancy); either way the next trip's indices
// and lengths are loaded before this trip's gathers are summed, so that the
// latency of the streamed matrix overlaps that of the gathers.
template <int F, int NB>
__device__ __forceinline__ void gather_sum4(const int* __restrict__ ind,
                                            const void* __restrict__ len,
                                            const float* __restrict__ X,
                                            long long e, long long e1,
                                            int lanes, float* acc)
{
  using Raw = typename Quad<F>::Raw;
  const long long step = 4LL * lanes;
  if (NB <= 2 && e + step < e1)
  {
    int4 i0 = *reinterpret_cast<const int4*>(ind + e);
    int4 i1 = *reinterpret_cast<const int4*>(ind + e + step);
    Raw q0 = Quad<F>::load(len, e);
    Raw q1 = Quad<F>::load(len, e + step);
    for (;;)
    {
      Vec<NB> xv[8];
      xv[0] = Vec<NB>::load(X, i0.x);
      xv[1] = Vec<NB>::load(X, i0.y);
      xv[2] = Vec<NB>::load(X, i0.z);
      xv[3] = Vec<NB>::load(X, i0.w);
      xv[4] = Vec<NB>::load(X, i1.x);
      xv[5] = Vec<NB>::load(X, i1.y);
      xv[6] = Vec<NB>::load(X, i1.z);
      xv[7] = Vec<NB>::load(X, i1.w);
      const long long en = e + 2 * step;
      const bool more = en + step < e1;
      int4 n0 = i0, n1 = i1;
      Raw p0 = q0, p1 = q1;
      if (more)
      {
        n0 = *reinterpret_cast<const int4*>(ind + en);
        n1 = *reinterpret_cast<const int4*>(ind + en + step);
        p0 = Quad<F>::load(len, en);
        p1 = Quad<F>::load(len, en + step);
      }
      float l[8];
      Quad<F>::dec(q0, l);
      Quad<F>::dec(q1, l + 4);
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = fmaf(l[u], xv[u].v[b], acc[b]);
      e = en;
      if (!more) break;
      i0 = n0;
      i1 = n1;
      q0 = p0;
      q1 = p1;
    }
  }
  if (NB > 2 && e < e1)
  {
    int4 i0 = *reinterpret_cast<const int4*>(ind + e);
    Raw q0 = Quad<F>::load(len, e);
    for (;;)
    {
      Vec<NB> xv[4];
      xv[0] = Vec<NB>::load(X, i0.x);
      xv[1] = Vec<NB>::load(X, i0.y);
      xv[2] = Vec<NB>::load(X, i0.z);
      xv[3] = Vec<NB>::load(X, i0.w);
      const long long en = e + step;
      const bool more = en < e1;
      int4 n0 = i0;
      Raw p0 = q0;
      if (more)
      {
        n0 = *reinterpret_cast<const int4*>(ind + en);
        p0 = Quad<F>::load(len, en);
      }
      float l[4];
      Quad<F>::dec(q0, l);
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = fmaf(l[u], xv[u].v[b], acc[b]);
      e = en;
      if (!more) break;
      i0 = n0;
      q0 = p0;
    }
    return;
  }
  for (; e < e1; e += step)
  {
    const int4 i0 = *reinterpret_cast<const int4*>(ind + e);
    float l[4];
    Quad<F>::dec(Quad<F>::load(len, e), l);
    Vec<NB> xv[4];
    xv[0] = Vec<NB>::load(X, i0.x);
    xv[1] = Vec<NB>::load(X, i0.y);
    xv[2] = Vec<NB>::load(X, i0.z);
    xv[3] = Vec<NB>::load(X, i0.w);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[b] = fmaf(l[u], xv[u].v[b], acc[b]);
  }
}

template <int NB>
__device__ __forceinline__ void lane_reduce(float* acc, int lanes)
{
  for (int o = lanes >> 1; o > 0; o >>= 1)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] += __shfl_xor(acc[b], o);
}

struct RowData
{
  const long long* off;
  const int* idx;
  const void* len;
  const int* bin;
  const float* factor;
  const unsigned char* yrow;
};

// Forward projection of the rows of one subset: sino[bin] = line integral.
template <int F>
__global__ __launch_bounds__(kBlock) void k_pc_forward(SubsetRows sr,
                                                       RowData rd, int lanes,
                                                       const float* __restrict__ img,
                                                       const float* __restrict__ img_y,
                                                       float* __restrict__ sino)
{
  int r, r0, k;
  const bool active = subset_row(sr, lanes, &r, &r0, &k);
  const int lane = (int)threadIdx.x & (lanes - 1);
  float acc[1] = {0.0f};
  if (active)
    gather_sum4<F, 1>(rd.idx, rd.len, rd.yrow[r] ? img_y : img,
                      rd.off[r] + 4LL * lane, rd.off[r + 1], lanes, acc);
  lane_reduce<1>(acc, lanes);
  if (active && lane == 0) sino[rd.bin[r]] = acc[0] * rd.factor[r];
}

struct OsemRowsIO
{
  const float* y[kMaxCacheBatch];
  const float* bias[kMaxCacheBatch];
  int n_images;
};

// OSEM line integrals of NB images (image b at X[v * NB + b]) over one
// subset's rows, and the per-row backprojection weights
// w[r * NB + b] = factor * y_b / (line_b + bias_b) when that line exceeds
// 1e-5, else 0 (reconAlgos.cc:126-158).
template <int F, int NB>
__global__ __launch_bounds__(kBlock) void k_pc_osem_rows(SubsetRows sr,
                                                         RowData rd, int lanes,
                                                         const float* __restrict__ X,
                                                         const float* __restrict__ XY_,
                                                         OsemRowsIO io,
                                                         float* __restrict__ w)
{
  int r, r0, k;
  const bool active = subset_row(sr, lanes, &r, &r0, &k);
  const int lane = (int)threadIdx.x & (lanes - 1);
  float acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = 0.0f;
  if (active)
    gather_sum4<F, NB>(rd.idx, rd.len, rd.yrow[r] ? XY_ : X,
                       rd.off[r] + 4LL * lane, rd.off[r + 1], lanes, acc);
  lane_reduce<NB>(acc, lanes);
  if (!active || lane != 0) return;
  const float f = rd.factor[r];
  const int bin = rd.bin[r];
#pragma unroll
  for (int b = 0; b < NB; ++b)
  {
    float out = 0.0f;
    if (b < io.n_images)
    {
      float line = acc[b] * f;
      if (io.bias[b]) line += io.bias[b][bin];
      if (line > kEpsF) out = io.y[b][bin] / line * f;
    }
    w[(long long)r * NB + b] = out;
  }
}

// Per-row backprojection weights of a sinogram (or of ones): the value
// times the row's length factor.
__global__ void k_pc_row_weights(const int* __restrict__ bin,
                                 const float* __restrict__ factor, int n_rows,
                                 const float* __restrict__ sino,
                                 float* __restrict__ w)
{
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       r < n_rows; r += (long long)gridDim.x * blockDim.x)
    w[r] = (sino ? sino[bin[r]] : 1.0f) * factor[r];
}

struct VoxOut
{
  float* p[kMaxCacheBatch];
};

// Backprojection of one subset: out_b[v] = sum over the subset's elements
// through voxel v of dec(len) * w[row * NB + b] (overwrites out_b).
template <int F, int NB>
__global__ __launch_bounds__(kBlock) void k_pc_backproject(
  const long long* __restrict__ voff, const int* __restrict__ trow,
  const void* __restrict__ tlen_, long long nvox, int lanes, int xcd_map,
  int zfast, const float* __restrict__ w, VoxOut out)
{
  long long v =
    ((long long)logical_block(blockIdx.x, gridDim.x, xcd_map) * blockDim.x +
     threadIdx.x) / lanes;
  // zfast = nz: work items walk the voxels z fastest, so that a wave's
  // voxels are z-neighbours (crossed by ring pairs of adjacent centre
  // slices, whose rows lie next to each other).
  if (zfast > 0 && v < nvox) v = (v % zfast) * (nvox / zfast) + v / zfast;
  const int lane = (int)threadIdx.x & (lanes - 1);
  const bool active = v < nvox;
  float acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = 0.0f;
  if (active)
    gather_sum4<F, NB>(trow, tlen_, w, voff[v] + 4LL * lane, voff[v + 1],
                       lanes, acc);
  lane_reduce<NB>(acc, lanes);
  if (!active || lane != 0) return;
#pragma unroll
  for (int b = 0; b < NB; ++b)
    if (out.p[b]) out.p[b][v] = acc[b];
}

struct ImgIn
{
  const float* p[kMaxCacheBatch];
};

// The gather copies of the images of one launch, written in their own
// order (so that each wave's stores are contiguous): blockIdx.y < nb_x
// fills row b = blockIdx.y of the x-fastest copy (a = x), the others the
// y-fastest copy (b = x, a = y); blockIdx.z is a group of zq slices; a
// thread writes one (a, z-slice, image) element, gathering it from the
// x-fastest images. Slices past nz (layout padding) get 0.
template <int NB>
__global__ void k_pc_pack(float* __restrict__ X, float* __restrict__ XT,
                          ImgIn in, int n_images, XY xy, int zq_shift, int nz)
{
  const bool xcopy = (int)blockIdx.y < xy.ny;
  if (xcopy && !X) return;
  const int b = xcopy ? (int)blockIdx.y : (int)blockIdx.y - xy.ny;
  const int na = xcopy ? xy.nx : xy.ny, nbb = xcopy ? xy.ny : xy.nx;
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= na * xy.zq * NB) return;
  const int img = t % NB;
  const int r = t / NB;
  const int zl = r & (xy.zq - 1);
  const int a = r >> zq_shift;
  const int zb = blockIdx.z;
  const int z = (zb << zq_shift) + zl;
  const int x = xcopy ? a : b, y = xcopy ? b : a;
  float v = 0.0f;
  if (z < nz && img < n_images)
    v = in.p[img][((long long)z * xy.ny + y) * xy.nx + x];
  const long long dst =
    ((((long long)zb * nbb + b) * na + a) << zq_shift) + zl;
  (xcopy ? X : XT)[dst * NB + img] = v;
}

// ---------------------------------------------------------------------------
// Sheet form of the path matrix
// ---------------------------------------------------------------------------
// With rows in tangential-major order, the rows of one (view, tangential
// bin) -- a family, every ring pair of one line of the xy-plane -- are
// contiguous, and all of them cross the same pixels of the xy-plane: the
// family's sheet, those pixels times all slices (at most nx + ny - 1 pixels,
// so 15 K voxels at 160 x 160 x 47). An element is then indexed within its
// family's sheet (pixel rank * nz + z, 16 bits), and a projector stages the
// sheet in LDS once per family instead of gathering each element from the
// image through the vector L1.

// Largest sheet, in floats, that a block stages (64 KB of LDS).
constexpr int kSheetMaxFloats = 16384;
constexpr int kSheetBlock = 1024;
constexpr int kSheetQuads = 2; // quads in flight per lane (2 measured best)
constexpr int kFamRowsMax = 2048; // rows of one family the fused pass holds

// First row of every family: the rows before the family's first LOR.
__global__ void k_sh_fam_rows(const int* __restrict__ rowid, int n_pairs,
                              int n_fam, int n_rows, int* __restrict__ fam_row)
{
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n_fam) fam_row[f] = rowid[(long long)f * n_pairs];
  else if (f == n_fam) fam_row[f] = n_rows;
}

// The voxels of up to kShK * 64 slots of each of two rows (lane ln takes
// slots ln, ln + 64, ...), all loads in flight together: v[h][k] is slot
// off[r_h] + ln + 64 k's voxel, or -1 (no row, past the row, or a
// zero-length element). Slots past kShK * 64 are left to the caller.
constexpr int kShK = 6;
__device__ __forceinline__ void sh_load_pair(
  const long long* __restrict__ off, const int* __restrict__ idx,
  const void* __restrict__ len, int lb,
  const unsigned char* __restrict__ yrow, const XY& xy, const int* r,
  int r_end, int ln, int (*v)[kShK], long long* lo, long long* hi,
  bool* yf)
{
#pragma unroll
  for (int h = 0; h < 2; ++h)
  {
    const bool ok = r[h] < r_end;
    lo[h] = ok ? off[r[h]] : 0;
    hi[h] = ok ? off[r[h] + 1] : 0;
    yf[h] = ok && yrow[r[h]] != 0;
  }
  int ix[2][kShK];
  unsigned lv[2][kShK];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int k = 0; k < kShK; ++k)
    {
      const long long e = lo[h] + ln + 64 * k;
      const bool in = e < hi[h];
      ix[h][k] = in ? idx[e] : 0;
      lv[h][k] = in ? (lb == 2 ? ((const unsigned short*)len)[e]
                               : ((const unsigned*)len)[e])
                    : 0u;
    }
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int k = 0; k < kShK; ++k)
      v[h][k] = lv[h][k] ? from_layout(xy, ix[h][k], yf[h]) : -1;
}

// Sheet pixels of every family, one block per family: a bitmap of the
// pixels its rows' contributing elements cross, built in LDS (a wave per
// row) and written out with the pixel count.
__global__ __launch_bounds__(kBlock) void k_sh_mark_fam(
  const int* __restrict__ fam_row, const long long* __restrict__ off,
  const int* __restrict__ idx, const void* __restrict__ len, int lb,
  const unsigned char* __restrict__ yrow, XY xy, int words,
  unsigned long long* __restrict__ bits, int* __restrict__ n_px)
{
  extern __shared__ unsigned long long wb[]; // [words]
  __shared__ int cnt;
  const int f = blockIdx.x;
  for (int i = threadIdx.x; i < words; i += kBlock) wb[i] = 0;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int nxy = xy.nx * xy.ny;
  const int wave = (int)threadIdx.x >> 6, ln = (int)threadIdx.x & 63;
  const int n_waves = kBlock / 64;
  const int rf = fam_row[f], rl = fam_row[f + 1];
  auto mark = [&](int v) {
    const int px = v % nxy;
    atomicOr(wb + px / 64, 1ULL << (px % 64));
  };
  for (int r0 = rf + wave; r0 < rl; r0 += 2 * n_waves)
  {
    const int r[2] = {r0, r0 + n_waves};
    int v[2][kShK];
    long long lo[2], hi[2];
    bool yf[2];
    sh_load_pair(off, idx, len, lb, yrow, xy, r, rl, ln, v, lo, hi, yf);
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
#pragma unroll
      for (int k = 0; k < kShK; ++k)
        if (v[h][k] >= 0) mark(v[h][k]);
      for (long long e = lo[h] + ln + 64 * kShK; e < hi[h]; e += 64)
      {
        const int w = tr_voxel(idx, len, lb, xy, yf[h], e, hi[h]);
        if (w >= 0) mark(w);
      }
    }
  }
  __syncthreads();
  int c = 0;
  for (int i = threadIdx.x; i < words; i += kBlock)
  {
    bits[(long long)f * words + i] = wb[i];
    c += __popcll(wb[i]);
  }
  atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) n_px[f] = cnt;
}

// Sheet pixels and sheet indices of every family, one block per family:
// the family's pixels from its bitmap (pixel order), ranked along its line
// when line_order is set (by the position of the pixel's centre along the
// family's line, from the first crystal to the second: every row of the
// family has the same two crystals in the xy-plane and walks from the
// first, Siddon.cc:112-412; a walk steps to a neighbouring pixel in the
// direction of the line, which moves the centre forward along it, so every
// row visits its pixels in rising rank, and each lane's elements rise too),
// written to fam_xy in rank order; then every slot of its rows gets its
// sheet index, rank * nz + z (0 for padding and zero-length elements),
// with the length's 16 bits above; list_quads[f] (cleared beforehand)
// gets the quads of the family's per-entry element lists, and fam_zr each
// sheet pixel's slices [lo, hi) that its elements cross (lo | hi << 16, in
// rank order; the slices outside hold no element of the family: the
// projectors neither stage nor sum them). Dynamic LDS: words * 12 bytes,
// 14 bytes per sheet pixel and 2 per sheet entry.
__global__ __launch_bounds__(kBlock) void k_sh_encode_fam(
  LorGeom lg, Grid g, int n_tang, const int* __restrict__ fam_row,
  const long long* __restrict__ off, const int* __restrict__ idx,
  const void* __restrict__ len, int lb,
  const unsigned char* __restrict__ yrow, XY xy, int nz, int words,
  const unsigned long long* __restrict__ bits,
  const int* __restrict__ fam_off, int line_order, int* __restrict__ fam_xy,
  unsigned* __restrict__ fam_zr, unsigned* __restrict__ sel,
  unsigned* __restrict__ list_quads)
{
  extern __shared__ unsigned long long wb[]; // [words]
  const int f = blockIdx.x;
  const int p0 = fam_off[f], L = fam_off[f + 1] - p0;
  int* wpre = reinterpret_cast<int*>(wb + words);  // [words]
  double* key = reinterpret_cast<double*>(wpre + ((words + 1) & ~1)); // [L]
  int* pxl = reinterpret_cast<int*>(key + L);       // [L]
  unsigned short* rank = reinterpret_cast<unsigned short*>(pxl + L); // [L]
  for (int i = threadIdx.x; i < words; i += kBlock)
    wb[i] = bits[(long long)f * words + i];
  __syncthreads();
  if (threadIdx.x == 0)
  {
    int c = 0;
    for (int k = 0; k < words; ++k)
    {
      wpre[k] = c;
      c += __popcll(wb[k]);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < words; k += kBlock)
  {
    unsigned long long b = wb[k];
    int o = wpre[k];
    while (b)
    {
      pxl[o++] = k * 64 + __ffsll((long long)b) - 1;
      b &= b - 1;
    }
  }
  __syncthreads();
  if (line_order)
  {
    const int view = f / n_tang, tt = f - view * n_tang;
    const int2 c = fir::view_crystals(lg, view, tt);
    const double2 a = lg.xy[c.x], b = lg.xy[c.y];
    const double dx = b.x - a.x, dy = b.y - a.y;
    for (int i = threadIdx.x; i < L; i += kBlock)
    {
      const int x = pxl[i] % g.row, y = pxl[i] / g.row;
      key[i] = (g.low[0] + (x + 0.5) * g.step[0] - a.x) * dx +
               (g.low[1] + (y + 0.5) * g.step[1] - a.y) * dy;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < L; i += kBlock)
    {
      const double ki = key[i];
      int k = 0;
      for (int j = 0; j < L; ++j)
        k += (key[j] < ki || (key[j] == ki && j < i)) ? 1 : 0;
      rank[i] = (unsigned short)k;
      fam_xy[p0 + k] = pxl[i];
    }
  }
  else
    for (int i = threadIdx.x; i < L; i += kBlock)
    {
      rank[i] = (unsigned short)i;
      fam_xy[p0 + i] = pxl[i];
    }
  __syncthreads();
  // elements per sheet entry, two 16-bit counts to a word (a family has
  // at most kFamRowsMax rows, each through an entry at most once)
  unsigned* ecnt = reinterpret_cast<unsigned*>(rank + ((L + 1) & ~1)); // [n_st / 2]
  const int n_st = L * nz;
  for (int i = threadIdx.x; i < (n_st + 1) / 2; i += kBlock) ecnt[i] = 0;
  __syncthreads();
  const int nxy = xy.nx * xy.ny;
  const int wave = (int)threadIdx.x >> 6, ln = (int)threadIdx.x & 63;
  const int n_waves = kBlock / 64;
  const int rf = fam_row[f], rl = fam_row[f + 1];
  const unsigned short* len16 = (const unsigned short*)len;
  auto encode = [&](int v, long long e) {
    unsigned out = 0;
    if (v >= 0)
    {
      const int px = v % nxy, z = v / nxy;
      const int k = px / 64;
      const int pr = wpre[k] + __popcll(wb[k] & ((1ULL << (px % 64)) - 1ULL));
      const unsigned en = (unsigned)(rank[pr] * nz + z);
      out = en | ((unsigned)len16[e] << 16);
      atomicAdd(ecnt + en / 2, 1u << (16 * (en & 1u)));
    }
    sel[e] = out;
  };
  for (int r0 = rf + wave; r0 < rl; r0 += 2 * n_waves)
  {
    const int r[2] = {r0, r0 + n_waves};
    int v[2][kShK];
    long long lo[2], hi[2];
    bool yf[2];
    sh_load_pair(off, idx, len, lb, yrow, xy, r, rl, ln, v, lo, hi, yf);
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
#pragma unroll
      for (int k = 0; k < kShK; ++k)
      {
        const long long e = lo[h] + ln + 64 * k;
        if (e < hi[h]) encode(v[h][k], e);
      }
      for (long long e = lo[h] + ln + 64 * kShK; e < hi[h]; e += 64)
        encode(tr_voxel(idx, len, lb, xy, yf[h], e, hi[h]), e);
    }
  }
  __syncthreads();
  auto count = [&](int en) { return (ecnt[en / 2] >> (16 * (en & 1))) & 0xffffu; };
  for (int p = threadIdx.x; p < L; p += kBlock)
  {
    int lo = nz, hi = 0;
    for (int z = 0; z < nz; ++z)
      if (count(p * nz + z))
      {
        lo = min(lo, z);
        hi = z + 1;
      }
    fam_zr[p0 + p] = lo < hi ? ((unsigned)lo | ((unsigned)hi << 16)) : 0u;
  }
  // the family's list quads (each entry's count rounded up to quads)
  unsigned q = 0;
  for (int i = threadIdx.x; i < n_st; i += kBlock)
    q += (((ecnt[i / 2] >> (16 * (i & 1))) & 0xffffu) + 3u) / 4u;
  for (int o = 32; o > 0; o >>= 1) q += __shfl_xor(q, o);
  if (ln == 0) atomicAdd(list_quads + f, q);
}

// Whether every lane's contributing elements (lanes = the quad layout's)
// have non-decreasing sheet pixels: clears *ok otherwise.
__global__ __launch_bounds__(kBlock) void k_sh_check_order(
  int n_rows, const long long* __restrict__ off,
  const unsigned* __restrict__ sel, int lanes, int nz, int* __restrict__ ok)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(t / lanes);
  if (r >= n_rows) return;
  const int j = (int)(t % lanes);
  int prev = 0;
  const long long end = off[r + 1];
  for (long long e = off[r] + 4LL * j; e < end; e += 4LL * lanes)
  {
    const uint4 q = *reinterpret_cast<const uint4*>(sel + e);
    const unsigned u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
      if ((u[k] >> 16) == 0) continue;
      const int p = (int)(u[k] & 0xffffu) / nz;
      if (p < prev)
      {
        atomicAnd(ok, 0);
        return;
      }
      prev = p;
    }
  }
}

// acc += sum over the elements of [e, e1) this lane owns of
// dec(len[i]) * sh[sidx[i]]: the quads and the order of gather_sum4's
// single-image path (so the sums are the same bit for bit), gathering
// from the staged sheet.
// The four 16-bit lengths of a quad of sheet elements, as Quad<F> loads
// them from the length array.
__device__ __forceinline__ uint2 sheet_lengths(uint4 q)
{
  return make_uint2((q.x >> 16) | (q.y & 0xffff0000u),
                    (q.z >> 16) | (q.w & 0xffff0000u));
}

template <int F>
__device__ __forceinline__ void sheet_sum(const unsigned* __restrict__ sel,
                                          const float* sh, long long e,
                                          long long e1, int lanes, float* acc)
{
  constexpr int Q = kSheetQuads;
  const long long step = 4LL * lanes;
  if (e + (Q - 1) * step < e1)
  {
    // Q quads per trip, the next Q loaded before this trip's sums
    uint4 q[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k)
      q[k] = *reinterpret_cast<const uint4*>(sel + e + k * step);
    for (;;)
    {
      float x[4 * Q];
#pragma unroll
      for (int k = 0; k < Q; ++k)
      {
        x[4 * k] = sh[q[k].x & 0xffffu];
        x[4 * k + 1] = sh[q[k].y & 0xffffu];
        x[4 * k + 2] = sh[q[k].z & 0xffffu];
        x[4 * k + 3] = sh[q[k].w & 0xffffu];
      }
      const long long en = e + Q * step;
      const bool more = en + (Q - 1) * step < e1;
      uint4 nq[Q];
#pragma unroll
      for (int k = 0; k < Q; ++k) nq[k] = q[k];
      if (more)
#pragma unroll
        for (int k = 0; k < Q; ++k)
          nq[k] = *reinterpret_cast<const uint4*>(sel + en + k * step);
      float l[4 * Q];
#pragma unroll
      for (int k = 0; k < Q; ++k) Quad<F>::dec(sheet_lengths(q[k]), l + 4 * k);
#pragma unroll
      for (int u = 0; u < 4 * Q; ++u) acc[0] = fmaf(l[u], x[u], acc[0]);
      e = en;
      if (!more) break;
#pragma unroll
      for (int k = 0; k < Q; ++k) q[k] = nq[k];
    }
  }
  for (; e < e1; e += step)
  {
    const uint4 q0 = *reinterpret_cast<const uint4*>(sel + e);
    float l[4];
    Quad<F>::dec(sheet_lengths(q0), l);
    acc[0] = fmaf(l[0], sh[q0.x & 0xffffu], acc[0]);
    acc[0] = fmaf(l[1], sh[q0.y & 0xffffu], acc[0]);
    acc[0] = fmaf(l[2], sh[q0.z & 0xffffu], acc[0]);
    acc[0] = fmaf(l[3], sh[q0.w & 0xffffu], acc[0]);
  }
}

// A family's sheet of one image staged into LDS from the image's
// pixel-major copy xpz[pixel * nz + z]: sh[p * nz + z] for the slices
// [lo, hi) of each sheet pixel p that the family's elements cross
// (fam_zr), consecutive threads reading consecutive slices of one pixel,
// eight loads in flight per thread. The entries outside are never read,
// except entry 0 by padding and zero-length slots (length 0), which is set
// to 0 when it lies outside.
__device__ __forceinline__ void stage_sheet(float* sh,
                                            const float* __restrict__ xpz,
                                            const int* __restrict__ fam_xy,
                                            const unsigned* __restrict__ fam_zr,
                                            int p0, int L, int nz)
{
  const int n_st = L * nz;
  if (threadIdx.x == 0)
  {
    const unsigned r = n_st ? fam_zr[p0] : 0u;
    if ((r & 0xffffu) != 0u || (r >> 16) == 0u) sh[0] = 0.0f;
  }
  for (int i0 = threadIdx.x; i0 < n_st; i0 += 8 * kSheetBlock)
  {
    float val[8];
    int at[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
    {
      const int i = i0 + u * kSheetBlock;
      at[u] = -1;
      val[u] = 0.0f;
      if (i < n_st)
      {
        const int p = i / nz, z = i - p * nz;
        const unsigned r = fam_zr[p0 + p];
        if ((unsigned)z >= (r & 0xffffu) && (unsigned)z < (r >> 16))
        {
          at[u] = i;
          val[u] = xpz[(long long)fam_xy[p0 + p] * nz + z];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (at[u] >= 0) sh[at[u]] = val[u];
  }
}

// k_pc_osem_rows for one image over the sheet form: one block per family
// of the subset (blockIdx.y: the subset's k-th view, blockIdx.x: the
// tangential bin) stages the family's sheet of `img` (x fastest) in LDS,
// then reduces its rows as k_pc_osem_rows does.
template <int F>
__global__ __launch_bounds__(kSheetBlock) void k_pc_osem_rows_sheet(
  const int* __restrict__ ford,
  const int* __restrict__ fam_row, const int* __restrict__ fam_off,
  const int* __restrict__ fam_xy, const unsigned* __restrict__ fam_zr,
  const unsigned* __restrict__ sel, RowData rd, int lanes,
  const float* __restrict__ img, int nxy, int nz, OsemRowsIO io,
  float* __restrict__ w)
{
  __shared__ float sh[kSheetMaxFloats];
  const int f = ford[blockIdx.x];
  const int r0 = fam_row[f], r1 = fam_row[f + 1];
  if (r0 == r1) return;
  const int p0 = fam_off[f];
  const int L = fam_off[f + 1] - p0;
  stage_sheet(sh, img, fam_xy, fam_zr, p0, L, nz);
  const int lane = (int)threadIdx.x & (lanes - 1);
  const int per_round = kSheetBlock / lanes;
  __syncthreads();
  for (int rb = r0; rb < r1; rb += per_round)
  {
    const int r = rb + (int)threadIdx.x / lanes;
    const bool active = r < r1;
    float acc[1] = {0.0f};
    if (active)
      sheet_sum<F>(sel, sh, rd.off[r] + 4LL * lane, rd.off[r + 1], lanes,
                   acc);
    lane_reduce<1>(acc, lanes);
    if (active && lane == 0)
    {
      const float fct = rd.factor[r];
      const int bin = rd.bin[r];
      float out = 0.0f;
      float line = acc[0] * fct;
      if (io.bias[0]) line += io.bias[0][bin];
      if (line > kEpsF) out = io.y[0][bin] / line * fct;
      w[r] = out;
    }
  }
}

// ---------------------------------------------------------------------------
// Batched line integrals over the sheet form, slab by slab
// ---------------------------------------------------------------------------
// The sheet of NB images (NB floats per sheet voxel) does not fit in LDS, so
// a family's block stages it in slabs of consecutive pixels (all slices).
// With the pixels ranked along the family's line, each lane meets its
// elements in rising pixel rank, so a slab's elements are, for every lane
// and row, a run that continues where the previous slab's stopped: a lane
// keeps one quad per row in hand, sums the elements that fall in the slab
// and moves on while the quad holds nothing beyond it. Each lane sums its
// elements in the order sheet_sum does (the zero-length ones, which add
// nothing there, are skipped), so every image's line integral is the
// one-image sheet pass's bit for bit.
constexpr int kSlabBlockMax = 1024;

// Pixel-major copies of NB images for the slab staging:
// X[(pixel * nz + z) * NB + b]. A block transposes a tile of kPzPx pixels
// by kPzZ slices of every image through LDS: its loads read runs of kPzPx
// pixels, its stores runs of kPzZ * NB floats per pixel.
constexpr int kPzPx = 64, kPzZ = 4;
constexpr int kPzPitch = kPzPx + 1;
template <int NB>
__device__ __forceinline__ void pack_pz_tile(float* __restrict__ X,
                                             const ImgIn& in, int n_images,
                                             int nxy, int nz, int bx, int bz,
                                             float* tile)
{
  constexpr int kPitch = kPzPitch;
  constexpr int kN = kPzPx * kPzZ * NB; // floats of the tile
  const int px0 = bx * kPzPx, z0 = bz * kPzZ;
  float v[kN / 256];
#pragma unroll
  for (int u = 0; u < kN / 256; ++u)
  {
    const int i = (int)threadIdx.x + 256 * u;
    const int p = i % kPzPx, r = i / kPzPx;
    const int b = r % NB, z = z0 + r / NB;
    v[u] = 0.0f;
    if (px0 + p < nxy && z < nz && b < n_images)
      v[u] = in.p[b][(long long)z * nxy + px0 + p];
  }
#pragma unroll
  for (int u = 0; u < kN / 256; ++u)
  {
    const int i = (int)threadIdx.x + 256 * u;
    tile[(i / kPzPx) * kPitch + i % kPzPx] = v[u];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kN / 256; ++u)
  {
    const int i = (int)threadIdx.x + 256 * u;
    const int b = i % NB, r = i / NB;
    const int zz = r % kPzZ, p = r / kPzZ;
    if (px0 + p < nxy && z0 + zz < nz)
      X[((long long)(px0 + p) * nz + z0 + zz) * NB + b] =
        tile[(zz * NB + b) * kPitch + p];
  }
}

template <int NB>
__global__ __launch_bounds__(256) void k_pc_pack_pz(float* __restrict__ X,
                                                    ImgIn in, int n_images,
                                                    int nxy, int nz)
{
  __shared__ float tile[kPzZ * NB * kPzPitch];
  pack_pz_tile<NB>(X, in, n_images, nxy, nz, blockIdx.x, blockIdx.y, tile);
}

template <int NB>
__device__ __forceinline__ void slab_fma(const float* __restrict__ xp, float l,
                                         float* acc)
{
  if constexpr (NB == 8)
  {
    const float4 a = reinterpret_cast<const float4*>(xp)[0];
    const float4 c = reinterpret_cast<const float4*>(xp)[1];
    acc[0] = fmaf(l, a.x, acc[0]);
    acc[1] = fmaf(l, a.y, acc[1]);
    acc[2] = fmaf(l, a.z, acc[2]);
    acc[3] = fmaf(l, a.w, acc[3]);
    acc[4] = fmaf(l, c.x, acc[4]);
    acc[5] = fmaf(l, c.y, acc[5]);
    acc[6] = fmaf(l, c.z, acc[6]);
    acc[7] = fmaf(l, c.w, acc[7]);
  }
  else
  {
    const float4 a = reinterpret_cast<const float4*>(xp)[0];
    acc[0] = fmaf(l, a.x, acc[0]);
    acc[1] = fmaf(l, a.y, acc[1]);
    acc[2] = fmaf(l, a.z, acc[2]);
    acc[3] = fmaf(l, a.w, acc[3]);
  }
}

// Stage slab `sb` (pixels [sb * slab_px, ...) of the family's sheet, all
// slices, NB floats per voxel: n4 float4 per pixel) into buf with direct
// global-to-LDS loads: a wave copies one pixel's run at a time, each
// instruction 64 consecutive float4 (lane j the j-th). Nothing lands in
// registers; the caller waits on the vector memory counter.
__device__ __forceinline__ void slab_stage(const float4* __restrict__ X4,
                                           const int* __restrict__ fam_xy,
                                           int p0, int L, int slab_px, int n4,
                                           int sb, float4* buf)
{
  const int pb = sb * slab_px;
  const int n = min(slab_px, L - pb);
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int ln = (int)threadIdx.x & 63;
  const int n_waves = (int)(blockDim.x >> 6);
  for (int p = wave; p < n; p += n_waves)
  {
    const float4* src = X4 + (long long)fam_xy[p0 + pb + p] * n4;
    float4* dst = buf + p * n4;
    for (int o = 0; o < n4; o += 64)
      if (o + ln < n4)
        __builtin_amdgcn_global_load_lds(src + o + ln, dst + o, 16, 0, 0);
  }
}

// k_pc_osem_rows for NB = 4 or 8 images over the sheet form: one block per
// family of the subset (blockIdx.y: the subset's k-th view, blockIdx.x: the
// tangential bin), blockDim.x = lanes * rows per round; thread t takes rows
// r0 + t / lanes + k * (blockDim.x / lanes), k < R (the host checks that R
// rounds cover the family), and keeps D quads of each in hand. X holds the
// images pixel-major (k_pc_pack_pz); slabs of slab_px pixels (all slices)
// are staged alternately into two LDS buffers of buf_f4 float4 each, the
// next slab's loads in flight while the current one is summed.
template <int F, int NB, int R, int D>
__device__ __forceinline__ void slab_family(
  int f, float4* slab, const int* __restrict__ fam_row,
  const int* __restrict__ fam_off, const int* __restrict__ fam_xy,
  const unsigned* __restrict__ sel, const RowData& rd, int lanes,
  const float* __restrict__ X, int nz, int slab_px, int buf_f4,
  const OsemRowsIO& io, float* __restrict__ w)
{
  const int r0 = fam_row[f], r1 = fam_row[f + 1];
  if (r0 == r1) return;
  const int p0 = fam_off[f];
  const int L = fam_off[f + 1] - p0;
  const int n4 = nz * NB / 4; // float4 per staged pixel
  const float4* X4 = reinterpret_cast<const float4*>(X);
  slab_stage(X4, fam_xy, p0, L, slab_px, n4, 0, slab);
  const int lane = (int)threadIdx.x & (lanes - 1);
  const int per_round = (int)blockDim.x / lanes;
  const long long step = 4LL * lanes;
  long long e[R], e1[R];
  uint4 q[R][D];
  float acc[R][NB];
#pragma unroll
  for (int k = 0; k < R; ++k)
  {
    const int r = r0 + (int)threadIdx.x / lanes + k * per_round;
    e[k] = 0;
    e1[k] = 0;
    if (r < r1)
    {
      e[k] = rd.off[r] + 4LL * lane;
      e1[k] = rd.off[r + 1];
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
      q[k][d] = e[k] + d * step < e1[k]
        ? *reinterpret_cast<const uint4*>(sel + e[k] + d * step)
        : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[k][b] = 0.0f;
  }
  const int n_slabs = (L + slab_px - 1) / slab_px;
  for (int sb = 0; sb < n_slabs; ++sb)
  {
    // this slab's loads have landed everywhere, and every thread is done
    // with the previous slab (whose buffer the next slab's loads overwrite)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (sb + 1 < n_slabs)
      slab_stage(X4, fam_xy, p0, L, slab_px, n4, sb + 1,
                 slab + ((sb + 1) & 1) * buf_f4);
    const int pb = sb * slab_px;
    const unsigned lo = (unsigned)(pb * nz);
    const unsigned hi = (unsigned)((pb + min(slab_px, L - pb)) * nz);
    const float* sl = reinterpret_cast<const float*>(slab + (sb & 1) * buf_f4);
    // a row whose quad holds elements beyond the slab is parked (its
    // elements in the slab are summed once) until the next slab
    bool park[R];
#pragma unroll
    for (int k = 0; k < R; ++k) park[k] = false;
    for (;;)
    {
      bool more = false;
#pragma unroll
      for (int k = 0; k < R; ++k)
      {
        if (e[k] < e1[k] && !park[k])
        {
          const uint4 qq = q[k][0];
          float l[4];
          Quad<F>::dec(sheet_lengths(qq), l);
          const unsigned u4[4] = {qq.x, qq.y, qq.z, qq.w};
          bool done = true;
#pragma unroll
          for (int u = 0; u < 4; ++u)
          {
            const unsigned si = u4[u] & 0xffffu;
            if ((u4[u] >> 16) != 0 && si >= lo)
            {
              if (si >= hi)
                done = false;
              else
                slab_fma<NB>(sl + (si - lo) * NB, l[u], acc[k]);
            }
          }
          park[k] = !done;
          if (done)
          {
            e[k] += step;
#pragma unroll
            for (int d = 0; d + 1 < D; ++d) q[k][d] = q[k][d + 1];
            const long long en = e[k] + (D - 1) * step;
            if (en < e1[k])
              q[k][D - 1] = *reinterpret_cast<const uint4*>(sel + en);
            more = more || e[k] < e1[k];
          }
        }
      }
      if (!more) break;
    }
  }
#pragma unroll
  for (int k = 0; k < R; ++k)
  {
    lane_reduce<NB>(acc[k], lanes);
    const int r = r0 + (int)threadIdx.x / lanes + k * per_round;
    if (r < r1 && lane == 0)
    {
      const float fc = rd.factor[r];
      const int bin = rd.bin[r];
#pragma unroll
      for (int b = 0; b < NB; ++b)
      {
        float out = 0.0f;
        if (b < io.n_images)
        {
          float line = acc[k][b] * fc;
          if (io.bias[b]) line += io.bias[b][bin];
          if (line > kEpsF) out = io.y[b][bin] / line * fc;
        }
        w[(long long)r * NB + b] = out;
      }
    }
  }
}

template <int F, int NB, int R, int D>
__global__ __launch_bounds__(kSlabBlockMax) void k_pc_osem_rows_slab(
  const int* __restrict__ ford,
  const int* __restrict__ fam_row, const int* __restrict__ fam_off,
  const int* __restrict__ fam_xy, const unsigned* __restrict__ sel,
  RowData rd, int lanes, const float* __restrict__ X, int nz, int slab_px,
  int buf_f4, OsemRowsIO io, float* __restrict__ w)
{
  extern __shared__ float4 slab[];
  slab_family<F, NB, R, D>(ford[blockIdx.x], slab, fam_row, fam_off, fam_xy,
                           sel, rd, lanes, X, nz, slab_px, buf_f4, io, w);
}

// ---------------------------------------------------------------------------
// Family backprojection over the sheet form
// ---------------------------------------------------------------------------
// The block that computed a family's line integrals also backprojects its
// weights, which it keeps in LDS: every sheet entry (pixel rank * nz + z)
// has a list of the family's elements through that voxel (row in family
// and length, rows ascending, padded to whole quads, stored as ListQuads),
// and the block sums each list against the LDS weights into a partial
// sheet (global entry (fam_off[f] + pixel rank) * nz + z). A reduction then sums,
// for every voxel, the partials of the subset's families whose sheet holds
// its pixel, in family order. No atomics: the sums are the same every run.

// Lists, count pass: the padded quad total of each family's lists.
constexpr int kFlBlock = 1024;
constexpr long long kFlStageQuads = 1LL << 26; // 1 GiB of list staging
constexpr int kFlRounds = 8; // list rounds a lane has in flight (4: 1-2% slower)

// A quad of list elements as stored: x = the first element's row in the
// family (11 bits, a family has at most kFamRowsMax rows) and the three
// row steps to the next elements (7 bits each; a list is in rising row
// order, and a padding element, length 0, repeats the previous row),
// y and z = the four 16-bit lengths as the length array holds them.
// 12 bytes instead of four (row | length << 16) words.
struct alignas(4) ListQuad
{
  unsigned x, y, z;
};
constexpr int kFlRowBits = 11;
constexpr int kFlStepBits = 7;
static_assert(kFamRowsMax <= (1 << kFlRowBits), "list rows exceed 11 bits");

__device__ __forceinline__ ListQuad list_quad(const unsigned* fl, long long q)
{
  return *reinterpret_cast<const ListQuad*>(fl + 3 * q);
}

__device__ __forceinline__ void list_rows(unsigned x, unsigned* r)
{
  constexpr unsigned m = (1u << kFlStepBits) - 1u;
  r[0] = x & ((1u << kFlRowBits) - 1u);
  r[1] = r[0] + ((x >> kFlRowBits) & m);
  r[2] = r[1] + ((x >> (kFlRowBits + kFlStepBits)) & m);
  r[3] = r[2] + (x >> (kFlRowBits + 2 * kFlStepBits));
}

__device__ __forceinline__ void fl_count(const int* __restrict__ fam_row,
                                         const long long* __restrict__ off,
                                         const unsigned* __restrict__ sel,
                                         int f, int n_st, unsigned* cnt)
{
  for (int i = threadIdx.x; i < n_st; i += kFlBlock) cnt[i] = 0;
  __syncthreads();
  const int r0 = fam_row[f], r1 = fam_row[f + 1];
  if (r0 < r1)
    for (long long e = off[r0] + threadIdx.x; e < off[r1]; e += kFlBlock)
    {
      const unsigned u = sel[e];
      if (u >> 16) atomicAdd(cnt + (u & 0xffffu), 1u);
    }
  __syncthreads();
}

// Exclusive scan of the padded quad counts of cnt[0 .. n_st) in place;
// returns the total to every thread.
__device__ __forceinline__ unsigned fl_scan(unsigned* cnt, int n_st,
                                            unsigned* part)
{
  const int chunk = (n_st + kFlBlock - 1) / kFlBlock;
  const int a = (int)threadIdx.x * chunk;
  const int b = min(n_st, a + chunk);
  unsigned t = 0;
  for (int i = a; i < b; ++i) t += (cnt[i] + 3u) / 4u;
  part[threadIdx.x] = t;
  __syncthreads();
  for (int d = 1; d < kFlBlock; d <<= 1) // inclusive Hillis-Steele scan
  {
    const unsigned v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned run = part[threadIdx.x] - t;
  for (int i = a; i < b; ++i)
  {
    const unsigned c = cnt[i];
    cnt[i] = run;
    run += (c + 3u) / 4u;
  }
  const unsigned total = part[kFlBlock - 1];
  __syncthreads();
  return total;
}

// Lists, fill pass: first quads of the family's entries (fl_off) and the
// elements (fl, cleared beforehand), each list in row order.
__global__ __launch_bounds__(kFlBlock) void k_fl_fill(
  const int* __restrict__ fam_row, const int* __restrict__ fam_off, int nz,
  const long long* __restrict__ off, const unsigned* __restrict__ sel,
  const long long* __restrict__ fbase, unsigned* __restrict__ fl_off,
  unsigned* __restrict__ fl)
{
  // cnt: the counts, then the list bases, then each group's row bits;
  // cur: elements placed so far per entry
  __shared__ unsigned cnt[kSheetMaxFloats];
  __shared__ unsigned short cur[kSheetMaxFloats];
  __shared__ unsigned part[kFlBlock];
  const int f = blockIdx.x;
  const int n_st = (fam_off[f + 1] - fam_off[f]) * nz;
  const long long g0 = (long long)fam_off[f] * nz;
  const long long base = fbase[f];
  fl_count(fam_row, off, sel, f, n_st, cnt);
  fl_scan(cnt, n_st, part);
  for (int i = threadIdx.x; i < n_st; i += kFlBlock)
  {
    fl_off[g0 + i] = (unsigned)(base + cnt[i]);
    cnt[i] = 0;
    cur[i] = 0;
  }
  __syncthreads();
  // Rows in groups of 32, each element placed at its entry's count so far
  // plus the group's earlier rows through that entry (a row crosses a voxel
  // once): every list in row order, no sorting. A wave per row.
  const int r0 = fam_row[f], r1 = fam_row[f + 1];
  const int wave = (int)threadIdx.x >> 6, ln = (int)threadIdx.x & 63;
  for (int ra = r0; ra < r1; ra += 32)
  {
    const int rb = min(ra + 32, r1);
    for (int r = ra + wave; r < rb; r += kFlBlock / 64)
      for (long long e = off[r] + ln; e < off[r + 1]; e += 64)
      {
        const unsigned u = sel[e];
        if (u >> 16) atomicOr(cnt + (u & 0xffffu), 1u << (r - ra));
      }
    __syncthreads();
    for (int r = ra + wave; r < rb; r += kFlBlock / 64)
      for (long long e = off[r] + ln; e < off[r + 1]; e += 64)
      {
        const unsigned u = sel[e];
        if (!(u >> 16)) continue;
        const unsigned en = u & 0xffffu;
        const unsigned k =
          cur[en] + (unsigned)__popc(cnt[en] & ((1u << (r - ra)) - 1u));
        fl[(unsigned long long)fl_off[g0 + en] * 4 + k] =
          (unsigned)(r - r0) | (u & 0xffff0000u);
      }
    __syncthreads();
    // the entry's first row in the group advances its count and clears
    // the bits
    for (int r = ra + wave; r < rb; r += kFlBlock / 64)
      for (long long e = off[r] + ln; e < off[r + 1]; e += 64)
      {
        const unsigned u = sel[e];
        if (!(u >> 16)) continue;
        const unsigned en = u & 0xffffu;
        const unsigned b = cnt[en];
        if ((b & ((1u << (r - ra)) - 1u)) == 0)
        {
          cur[en] = (unsigned short)(cur[en] + __popc(b));
          cnt[en] = 0;
        }
      }
    __syncthreads();
  }
}

// The sheet entries (fam_off positions) of every pixel, per subset: counts.
__global__ void k_fr_count(const int* __restrict__ fam_off,
                           const int* __restrict__ fam_xy, int n_fam,
                           int total, int n_tang, int S, int nxy,
                           unsigned* __restrict__ cnt)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int lo = 0, hi = n_fam - 1; // family of entry i
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (fam_off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const int s = (lo / n_tang) % S;
  atomicAdd(cnt + (long long)s * (nxy + 1) + fam_xy[i], 1u);
}

// ... and the entries themselves, as positions among the sheet pixels of
// their subset's families (fam_loc: a family's first position), which is
// where the launches over that subset keep the family's partial sheet,
// with the slice range of each (the partials outside are never written).
__global__ void k_fr_fill(const int* __restrict__ fam_off,
                          const int* __restrict__ fam_xy,
                          const unsigned* __restrict__ fam_zr,
                          const int* __restrict__ fam_loc, int n_fam,
                          int total, int n_tang, int S, int nxy,
                          const long long* __restrict__ off,
                          unsigned* __restrict__ cursor, int2* __restrict__ ent)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int lo = 0, hi = n_fam - 1;
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (fam_off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const int s = (lo / n_tang) % S;
  const long long k = (long long)s * (nxy + 1) + fam_xy[i];
  const long long at = off[k] + atomicAdd(cursor + k, 1u);
  ent[at] = make_int2(fam_loc[lo] + (i - fam_off[lo]), (int)fam_zr[i]);
}

// Each pixel's entries in increasing order (family order), their slice
// ranges alongside.
__global__ void k_fr_sort(const long long* __restrict__ off, long long n,
                          int2* __restrict__ ent)
{
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const long long a = off[k], b = off[k + 1];
  for (long long i = a + 1; i < b; ++i)
  {
    const int2 e = ent[i];
    long long j = i - 1;
    while (j >= a && ent[j].x > e.x)
    {
      ent[j + 1] = ent[j];
      --j;
    }
    ent[j + 1] = e;
  }
}

// Lists, interleave pass: the family's entries in slots sorted by list
// length (longest first), 64 slots to a group (one wave), and each group's
// lists stored round by round: round m holds the m-th quad of every slot
// whose list is longer than m, in slot order, so that a wave's loads of one
// round are contiguous. Per slot: entry << 16 | list quads, and the group's
// first quad. Families f_first + blockIdx.x; their lists are written to
// fl[4 * (quad - dst_base)] (a staging buffer for a run of families).
constexpr int kFlMaxQuads = kFamRowsMax / 4; // a voxel is on a row at most once

// Sort key of a family's entry e (pixel rank * nz + z) holding q list
// quads, slots taken in decreasing key: entries with a list by length, then
// the empty ones inside their pixel's slice range (whose partials are
// written as 0, as the per-voxel sums read them), then the empty ones
// outside it (never read or written: the slots past the family's count of
// active slots).
__device__ __forceinline__ unsigned fl_sort_key(unsigned q,
                                                const unsigned* zr, int e,
                                                int nz)
{
  if (q) return min(q, (unsigned)kFlMaxQuads) + 1u;
  const int p = e / nz, z = e - p * nz;
  const unsigned r = zr[p];
  return (unsigned)z >= (r & 0xffffu) && (unsigned)z < (r >> 16) ? 1u : 0u;
}

__global__ __launch_bounds__(kFlBlock) void k_fl_interleave(
  int f_first, const int* __restrict__ fam_off, int nz,
  const unsigned* __restrict__ fam_zr, const long long* __restrict__ fbase,
  const unsigned* __restrict__ cl_off, const unsigned* __restrict__ cl,
  unsigned* __restrict__ fl_slot, unsigned* __restrict__ fl_gb,
  unsigned* __restrict__ fl, int* __restrict__ fl_nact, long long dst_base)
{
  __shared__ unsigned short lq[kSheetMaxFloats];
  __shared__ unsigned short ord[kSheetMaxFloats];
  __shared__ unsigned hist[kFlMaxQuads + 2];
  __shared__ unsigned gpre[kSheetMaxFloats / 64 + 1];
  const int f = f_first + (int)blockIdx.x;
  const int n_st = (fam_off[f + 1] - fam_off[f]) * nz;
  if (n_st == 0) return;
  const long long g0 = (long long)fam_off[f] * nz;
  const long long base = fbase[f];
  const int n_groups = (n_st + 63) / 64;
  const unsigned* zr = fam_zr + fam_off[f];
  for (int i = threadIdx.x; i <= kFlMaxQuads + 1; i += kFlBlock) hist[i] = 0;
  for (int i = threadIdx.x; i <= n_groups; i += kFlBlock) gpre[i] = 0;
  __syncthreads();
  for (int e = threadIdx.x; e < n_st; e += kFlBlock)
  {
    const unsigned q = cl_off[g0 + e + 1] - cl_off[g0 + e];
    lq[e] = (unsigned short)q;
    atomicAdd(hist + fl_sort_key(q, zr, e, nz), 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0)
  {
    unsigned run = 0;
    for (int v = kFlMaxQuads + 1; v >= 0; --v)
    {
      const unsigned c = hist[v];
      hist[v] = run;
      run += c;
    }
    fl_nact[f] = (int)hist[0];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n_st; e += kFlBlock)
    ord[atomicAdd(hist + fl_sort_key(lq[e], zr, e, nz), 1u)] =
      (unsigned short)e;
  __syncthreads();
  for (int j = threadIdx.x; j < n_st; j += kFlBlock)
    atomicAdd(gpre + j / 64, (unsigned)lq[ord[j]]);
  __syncthreads();
  if (threadIdx.x == 0)
  {
    unsigned run = 0;
    for (int g = 0; g < n_groups; ++g)
    {
      const unsigned c = gpre[g];
      gpre[g] = run;
      run += c;
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n_st; j += kFlBlock)
  {
    const int e = ord[j];
    fl_slot[g0 + j] = ((unsigned)e << 16) | lq[e];
    fl_gb[g0 + j] = (unsigned)(base + gpre[j / 64]);
  }
  const int wave = (int)threadIdx.x >> 6, ln = (int)threadIdx.x & 63;
  const unsigned long long below = (1ull << ln) - 1ull;
  for (int g = wave; g < n_groups; g += kFlBlock / 64)
  {
    const int j = g * 64 + ln;
    const int e = j < n_st ? ord[j] : 0;
    const int q = j < n_st ? lq[e] : 0;
    const unsigned src = j < n_st ? cl_off[g0 + e] : 0u;
    const long long gb = base + gpre[g];
    unsigned roff = 0;
    for (int m = 0;; ++m)
    {
      const unsigned long long act = __ballot(m < q);
      if (act == 0) break;
      if (m < q)
      {
        const long long d = gb + roff + __popcll(act & below) - dst_base;
        *reinterpret_cast<uint4*>(fl + 4 * d) =
          *reinterpret_cast<const uint4*>(cl + 4LL * (src + m));
      }
      roff += (unsigned)__popcll(act);
    }
  }
}

// The family lists in their final layout in one pass per family (the
// interleave pass's layout, without the contiguous lists and the staging
// buffer): the entries' list lengths, slots ordered by length (longest
// first), each 64-slot group's round offsets (round m of a group holds the
// m-th quad of its slots longer than m, which are a prefix of the group),
// then the elements placed as k_fl_fill places them (rows in groups of 32,
// each list in row order) straight at their final positions, and each
// list's last quad padded with zeros. The families f_first + blockIdx.x
// write their quads [q, q + n) as four words each at fl[4 * (q - dst_base)]
// (a staging buffer that k_fl_pack then stores as ListQuads). A family
// whose round-offset table exceeds kFlRTable entries sets *overflow and
// writes nothing (its lists then take the two-pass route).
constexpr int kFlRTable = 4096;
constexpr int kFlK = 6; // elements per lane of a row held in registers

__global__ __launch_bounds__(kFlBlock) void k_fl_build(
  int f_first, const int* __restrict__ fam_row,
  const int* __restrict__ fam_off, int nz,
  const unsigned* __restrict__ fam_zr, const long long* __restrict__ off,
  const unsigned* __restrict__ sel, const long long* __restrict__ fbase,
  unsigned* __restrict__ fl_slot, unsigned* __restrict__ fl_gb,
  unsigned* __restrict__ fl, long long dst_base, int* __restrict__ fl_nact,
  int* __restrict__ overflow)
{
  __shared__ unsigned cnt[kSheetMaxFloats];          // counts, then row bits
  __shared__ unsigned short cur[kSheetMaxFloats];    // quads per slot, then
                                                     // elements placed
  __shared__ unsigned short slot[kSheetMaxFloats];   // entry -> slot
  __shared__ unsigned rtab[kFlRTable];
  __shared__ unsigned hist[kFlMaxQuads + 2];
  __shared__ unsigned gbase[kSheetMaxFloats / 64 + 1];
  __shared__ unsigned rofs[kSheetMaxFloats / 64 + 1];
  __shared__ int too_big;
  const int f = f_first + (int)blockIdx.x;
  const int n_st = (fam_off[f + 1] - fam_off[f]) * nz;
  if (n_st == 0) return;
  const long long g0 = (long long)fam_off[f] * nz;
  const long long base = fbase[f];
  const int n_groups = (n_st + 63) / 64;
  const int r0 = fam_row[f], r1 = fam_row[f + 1];
  for (int i = threadIdx.x; i < n_st; i += kFlBlock) cnt[i] = 0;
  for (int i = threadIdx.x; i <= kFlMaxQuads + 1; i += kFlBlock) hist[i] = 0;
  if (threadIdx.x == 0) too_big = 0;
  __syncthreads();
  if (r0 < r1)
  {
    // eight loads in flight per thread
    const long long ea = off[r0], eb = off[r1];
    for (long long e0 = ea + threadIdx.x; e0 < eb; e0 += 8 * kFlBlock)
    {
      unsigned u[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
      {
        const long long e = e0 + (long long)k * kFlBlock;
        u[k] = e < eb ? sel[e] : 0u;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (u[k] >> 16) atomicAdd(cnt + (u[k] & 0xffffu), 1u);
    }
  }
  __syncthreads();
  const unsigned* zr = fam_zr + fam_off[f];
  for (int e = threadIdx.x; e < n_st; e += kFlBlock)
    atomicAdd(hist + fl_sort_key((cnt[e] + 3u) / 4u, zr, e, nz), 1u);
  __syncthreads();
  if (threadIdx.x == 0)
  {
    unsigned run = 0;
    for (int v = kFlMaxQuads + 1; v >= 0; --v)
    {
      const unsigned c = hist[v];
      hist[v] = run;
      run += c;
    }
    fl_nact[f] = (int)hist[0];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n_st; e += kFlBlock)
  {
    const unsigned q = (cnt[e] + 3u) / 4u;
    const unsigned sl = atomicAdd(hist + fl_sort_key(q, zr, e, nz), 1u);
    slot[e] = (unsigned short)sl;
    cur[sl] = (unsigned short)q;
  }
  __syncthreads();
  if (threadIdx.x == 0)
  {
    unsigned o = 0;
    for (int g = 0; g < n_groups; ++g)
    {
      rofs[g] = o;
      o += cur[64 * g] + 1u; // the group's longest list comes first
    }
    too_big = o > (unsigned)kFlRTable;
  }
  __syncthreads();
  if (too_big)
  {
    if (threadIdx.x == 0) atomicOr(overflow, 1);
    return;
  }
  for (int g = threadIdx.x; g < n_groups; g += kFlBlock)
  {
    const int s0 = 64 * g;
    const int M = cur[s0];
    int i = min(64, n_st - s0); // slots longer than m: a prefix of the group
    unsigned R = 0;
    for (int m = 0; m < M; ++m)
    {
      while (i > 0 && cur[s0 + i - 1] <= m) --i;
      rtab[rofs[g] + m] = R;
      R += (unsigned)i;
    }
    rtab[rofs[g] + M] = R;
  }
  __syncthreads();
  if (threadIdx.x == 0)
  {
    unsigned o = 0;
    for (int g = 0; g < n_groups; ++g)
    {
      gbase[g] = o;
      o += rtab[rofs[g] + cur[64 * g]];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n_st; e += kFlBlock)
  {
    const unsigned sl = slot[e];
    fl_slot[g0 + sl] = ((unsigned)e << 16) | cur[sl];
    fl_gb[g0 + sl] = (unsigned)(base + gbase[sl >> 6]);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_st; i += kFlBlock)
  {
    cnt[i] = 0;
    cur[i] = 0;
  }
  __syncthreads();
  // Rows in groups of 32, a wave taking two rows of each: the rows' first
  // 64 * kFlK elements are loaded into registers once (all in flight
  // together) and serve the group's three passes; longer rows' remainders
  // are read again in each pass.
  const int wave = (int)threadIdx.x >> 6, ln = (int)threadIdx.x & 63;
  for (int ra = r0; ra < r1; ra += 32)
  {
    const int rb = min(ra + 32, r1);
    unsigned v[2][kFlK];
    long long lo[2], hi[2];
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
      const int r = ra + wave + 16 * h;
      lo[h] = r < rb ? off[r] : 0;
      hi[h] = r < rb ? off[r + 1] : 0;
#pragma unroll
      for (int k = 0; k < kFlK; ++k)
      {
        const long long e = lo[h] + ln + 64 * k;
        v[h][k] = e < hi[h] ? sel[e] : 0u;
      }
    }
    // pass 1: the group's rows through each entry
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
      const unsigned bit = 1u << (wave + 16 * h);
#pragma unroll
      for (int k = 0; k < kFlK; ++k)
        if (v[h][k] >> 16) atomicOr(cnt + (v[h][k] & 0xffffu), bit);
      for (long long e = lo[h] + ln + 64 * kFlK; e < hi[h]; e += 64)
      {
        const unsigned u = sel[e];
        if (u >> 16) atomicOr(cnt + (u & 0xffffu), bit);
      }
    }
    __syncthreads();
    // pass 2: each element at its entry's count plus the group's earlier
    // rows through it
    auto place = [&](unsigned u, int rl) {
      const unsigned en = u & 0xffffu;
      const unsigned c =
        cur[en] + (unsigned)__popc(cnt[en] & ((1u << rl) - 1u));
      const unsigned sl = slot[en], g = sl >> 6;
      const long long quad =
        base - dst_base + gbase[g] + rtab[rofs[g] + (c >> 2)] + (sl & 63u);
      fl[quad * 4 + (c & 3u)] = (unsigned)(ra + rl - r0) | (u & 0xffff0000u);
    };
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
      const int rl = wave + 16 * h;
#pragma unroll
      for (int k = 0; k < kFlK; ++k)
        if (v[h][k] >> 16) place(v[h][k], rl);
      for (long long e = lo[h] + ln + 64 * kFlK; e < hi[h]; e += 64)
      {
        const unsigned u = sel[e];
        if (u >> 16) place(u, rl);
      }
    }
    __syncthreads();
    // pass 3: the entry's first row in the group advances its count and
    // clears the bits
    auto advance = [&](unsigned u, int rl) {
      const unsigned en = u & 0xffffu;
      const unsigned b = cnt[en];
      if ((b & ((1u << rl) - 1u)) == 0)
      {
        cur[en] = (unsigned short)(cur[en] + __popc(b));
        cnt[en] = 0;
      }
    };
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
      const int rl = wave + 16 * h;
#pragma unroll
      for (int k = 0; k < kFlK; ++k)
        if (v[h][k] >> 16) advance(v[h][k], rl);
      for (long long e = lo[h] + ln + 64 * kFlK; e < hi[h]; e += 64)
      {
        const unsigned u = sel[e];
        if (u >> 16) advance(u, rl);
      }
    }
    __syncthreads();
  }
  // the zero padding of each list's last quad (cur: the lists' lengths)
  for (int en = threadIdx.x; en < n_st; en += kFlBlock)
  {
    const unsigned sl = slot[en], g = sl >> 6;
    for (unsigned c = cur[en]; c & 3u; ++c)
    {
      const long long quad =
        base - dst_base + gbase[g] + rtab[rofs[g] + (c >> 2)] + (sl & 63u);
      fl[quad * 4 + (c & 3u)] = 0u;
    }
  }
}

// The n list quads of a staging buffer (four row | length << 16 words
// each) as ListQuads at dst. Sets *overflow when a row step of some quad
// exceeds what the 7-bit fields hold, or the rows fall (the lists then
// cannot be stored).
__global__ __launch_bounds__(kBlock) void k_fl_pack(
  const unsigned* __restrict__ src, long long n, unsigned* __restrict__ dst,
  int* __restrict__ overflow)
{
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint4 v = reinterpret_cast<const uint4*>(src)[i];
  const unsigned u[4] = {v.x, v.y, v.z, v.w};
  // a quad starts with an element (padding only ends a list)
  unsigned row = u[0] & 0xffffu, x = row;
  bool ok = row < (1u << kFlRowBits);
#pragma unroll
  for (int k = 1; k < 4; ++k)
  {
    const unsigned r = (u[k] >> 16) ? (u[k] & 0xffffu) : row;
    const unsigned d = r - row; // wraps above the 7 bits if the rows fall
    ok = ok && d < (1u << kFlStepBits);
    x |= d << (kFlRowBits + (k - 1) * kFlStepBits);
    row = r;
  }
  if (!ok) atomicOr(overflow, 1);
  unsigned* o = dst + 3 * i;
  o[0] = x;
  o[1] = (u[0] >> 16) | (u[1] & 0xffff0000u);
  o[2] = (u[2] >> 16) | (u[3] & 0xffff0000u);
}

// k_pc_osem_rows_sheet plus the family's backprojection: the weights stay
// in LDS, each lane sums one sheet voxel's list against them, and the sums
// go to the family's partial sheet in `part`.
template <int F>
__global__ __launch_bounds__(kSheetBlock) void k_pc_osem_fam(
  const int* __restrict__ ford,
  const int* __restrict__ fam_row, const int* __restrict__ fam_off,
  const int* __restrict__ fam_xy, const unsigned* __restrict__ fam_zr,
  const unsigned* __restrict__ sel, RowData rd, int lanes,
  const float* __restrict__ img, int nxy, int nz, OsemRowsIO io,
  float* __restrict__ w, const unsigned* __restrict__ fl_slot,
  const unsigned* __restrict__ fl_gb, const unsigned* __restrict__ fl,
  const int* __restrict__ fl_nact, const int* __restrict__ fam_loc,
  float* __restrict__ part)
{
  __shared__ float sh[kSheetMaxFloats];
  __shared__ float wr[kFamRowsMax];
  const int f = ford[blockIdx.x];
  const int r0 = fam_row[f], r1 = fam_row[f + 1];
  const int p0 = fam_off[f];
  const int L = fam_off[f + 1] - p0;
  const int n_st = L * nz;
  float* pf = part + (long long)fam_loc[f] * nz;
  if (r0 == r1 || r1 - r0 > kFamRowsMax)
  {
    // no rows: the family's partials are 0 (a family never exceeds
    // kFamRowsMax rows: the host checks)
    for (int i = threadIdx.x; i < n_st; i += kSheetBlock) pf[i] = 0.0f;
    return;
  }
  stage_sheet(sh, img, fam_xy, fam_zr, p0, L, nz);
  const int lane = (int)threadIdx.x & (lanes - 1);
  const int per_round = kSheetBlock / lanes;
  __syncthreads();
  for (int rb = r0; rb < r1; rb += per_round)
  {
    const int r = rb + (int)threadIdx.x / lanes;
    const bool active = r < r1;
    float acc[1] = {0.0f};
    if (active)
      sheet_sum<F>(sel, sh, rd.off[r] + 4LL * lane, rd.off[r + 1], lanes, acc);
    lane_reduce<1>(acc, lanes);
    if (active && lane == 0)
    {
      const float fct = rd.factor[r];
      const int bin = rd.bin[r];
      float out = 0.0f;
      float line = acc[0] * fct;
      if (io.bias[0]) line += io.bias[0][bin];
      if (line > kEpsF) out = io.y[0][bin] / line * fct;
      if (w) w[r] = out;
      wr[r - r0] = out;
    }
  }
  __syncthreads();
  // each sheet entry's list against the weights in LDS: a wave takes a
  // group of 64 slots, a lane one entry, summed in row order (the same
  // sums every run), kFlRounds rounds of loads in flight
  const long long g0 = (long long)p0 * nz;
  const int wave = (int)threadIdx.x >> 6, ln = (int)threadIdx.x & 63;
  const unsigned long long below = (1ull << ln) - 1ull;
  // the slots with a list or inside their pixel's slice range come first
  const int n_act = fl_nact[f];
  for (int g = wave; g * 64 < n_act; g += kSheetBlock / 64)
  {
    const int j = g * 64 + ln;
    const unsigned sl = j < n_act ? fl_slot[g0 + j] : 0u;
    // every slot of a group holds the group's first quad: one
    // wave-uniform load instead of one per lane
    const unsigned gb =
      fl_gb[g0 + 64LL * __builtin_amdgcn_readfirstlane(g)];
    const int q = (int)(sl & 0xffffu);
    float acc = 0.0f;
    unsigned roff = 0;
    for (int m = 0;; m += kFlRounds)
    {
      unsigned long long act[kFlRounds];
#pragma unroll
      for (int k = 0; k < kFlRounds; ++k) act[k] = __ballot(m + k < q);
      if (act[0] == 0) break;
      ListQuad v[kFlRounds];
      unsigned o = roff;
#pragma unroll
      for (int k = 0; k < kFlRounds; ++k)
      {
        v[k] = m + k < q ? list_quad(fl, (long long)gb + o +
                                            __popcll(act[k] & below))
                         : ListQuad{0u, 0u, 0u};
        o += (unsigned)__popcll(act[k]);
      }
      roff = o;
#pragma unroll
      for (int k = 0; k < kFlRounds; ++k)
      {
        float l[4];
        unsigned rw[4];
        Quad<F>::dec(make_uint2(v[k].y, v[k].z), l);
        list_rows(v[k].x, rw);
        acc = fmaf(l[0], wr[rw[0]], acc);
        acc = fmaf(l[1], wr[rw[1]], acc);
        acc = fmaf(l[2], wr[rw[2]], acc);
        acc = fmaf(l[3], wr[rw[3]], acc);
      }
    }
    // into the sheet's LDS buffer (free once the line integrals are done),
    // so that the partial sheet goes out in contiguous runs below
    if (j < n_act) sh[sl >> 16] = acc;
  }
  __syncthreads();
  // the entries inside their pixel's slice range, which the per-voxel sums
  // read (the listed slots and the empty ones in range, written 0 above)
  for (int i = threadIdx.x; i < n_st; i += kSheetBlock)
  {
    const int p = i / nz, z = i - p * nz;
    const unsigned r = fam_zr[p0 + p];
    if ((unsigned)z >= (r & 0xffffu) && (unsigned)z < (r >> 16)) pf[i] = sh[i];
  }
}

// Family backprojection of NB images with given per-row weights
// w[r * NB + b] (a batch's line integrals, or a sinogram times the length
// factors): one block per family of the subset copies the family's weights
// to LDS and sums every sheet entry's list against them, as the second
// half of k_pc_osem_fam does for one image, into the partial sheets
// part[(entry * NB + b)], entry = fam_loc[f] * nz + sheet index.
template <int F, int NB>
__device__ __forceinline__ void fam_bp_family(
  int f, float* wr, const int* __restrict__ fam_row,
  const int* __restrict__ fam_off, const int* __restrict__ fam_loc, int nz,
  const float* __restrict__ w, const unsigned* __restrict__ fl_slot,
  const unsigned* __restrict__ fl_gb, const unsigned* __restrict__ fl,
  const int* __restrict__ fl_nact, float* __restrict__ part,
  float* stage = nullptr, const unsigned* __restrict__ fam_zr = nullptr)
{
  const int r0 = fam_row[f], r1 = fam_row[f + 1];
  const int p0 = fam_off[f];
  const int n_st = (fam_off[f + 1] - p0) * nz;
  float* pf = part + (long long)fam_loc[f] * nz * NB;
  if (r0 == r1 || r1 - r0 > kFamRowsMax)
  {
    for (int i = threadIdx.x; i < n_st * NB; i += kSheetBlock) pf[i] = 0.0f;
    return;
  }
  const int nw = (r1 - r0) * NB;
  for (int i = threadIdx.x; i < nw; i += kSheetBlock)
    wr[i] = w[(long long)r0 * NB + i];
  __syncthreads();
  const long long g0 = (long long)p0 * nz;
  const int wave = (int)threadIdx.x >> 6, ln = (int)threadIdx.x & 63;
  const unsigned long long below = (1ull << ln) - 1ull;
  // the slots with a list or inside their pixel's slice range come first
  const int n_act = fl_nact[f];
  for (int g = wave; g * 64 < n_act; g += kSheetBlock / 64)
  {
    const int j = g * 64 + ln;
    const unsigned sl = j < n_act ? fl_slot[g0 + j] : 0u;
    // every slot of a group holds the group's first quad: one
    // wave-uniform load instead of one per lane
    const unsigned gb =
      fl_gb[g0 + 64LL * __builtin_amdgcn_readfirstlane(g)];
    const int q = (int)(sl & 0xffffu);
    float acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = 0.0f;
    unsigned roff = 0;
    for (int m = 0;; m += kFlRounds)
    {
      unsigned long long act[kFlRounds];
#pragma unroll
      for (int k = 0; k < kFlRounds; ++k) act[k] = __ballot(m + k < q);
      if (act[0] == 0) break;
      ListQuad v[kFlRounds];
      unsigned o = roff;
#pragma unroll
      for (int k = 0; k < kFlRounds; ++k)
      {
        v[k] = m + k < q ? list_quad(fl, (long long)gb + o +
                                            __popcll(act[k] & below))
                         : ListQuad{0u, 0u, 0u};
        o += (unsigned)__popcll(act[k]);
      }
      roff = o;
#pragma unroll
      for (int k = 0; k < kFlRounds; ++k)
      {
        float l[4];
        unsigned rw[4];
        Quad<F>::dec(make_uint2(v[k].y, v[k].z), l);
        list_rows(v[k].x, rw);
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
          const float* wp = wr + rw[u] * NB;
#pragma unroll
          for (int b = 0; b < NB; ++b) acc[b] = fmaf(l[u], wp[b], acc[b]);
        }
      }
    }
    if (j < n_act)
    {
      if (stage)
        stage[sl >> 16] = acc[0];
      else
      {
        float* o = pf + (long long)(sl >> 16) * NB;
#pragma unroll
        for (int b = 0; b < NB; ++b) o[b] = acc[b];
      }
    }
  }
  if (stage)
  {
    // one image: the partial sheet gathered in LDS goes out in contiguous
    // runs of each pixel's slice range (as k_pc_osem_fam stores it)
    __syncthreads();
    for (int i = threadIdx.x; i < n_st; i += kSheetBlock)
    {
      const int p = i / nz, z = i - p * nz;
      const unsigned r = fam_zr[p0 + p];
      if ((unsigned)z >= (r & 0xffffu) && (unsigned)z < (r >> 16))
        pf[i] = stage[i];
    }
  }
}

template <int F, int NB>
__global__ __launch_bounds__(kSheetBlock) void k_pc_fam_bp(
  const int* __restrict__ ford,
  const int* __restrict__ fam_row, const int* __restrict__ fam_off,
  const int* __restrict__ fam_loc, int nz, const float* __restrict__ w,
  const unsigned* __restrict__ fl_slot, const unsigned* __restrict__ fl_gb,
  const unsigned* __restrict__ fl, const int* __restrict__ fl_nact,
  float* __restrict__ part, const unsigned* __restrict__ fam_zr)
{
  // one image: room for the partial sheet as well
  __shared__ float wr[kFamRowsMax * NB + (NB == 1 ? kSheetMaxFloats : 0)];
  fam_bp_family<F, NB>(ford[blockIdx.x], wr, fam_row, fam_off, fam_loc, nz, w,
                       fl_slot, fl_gb, fl, fl_nact, part,
                       NB == 1 ? wr + kFamRowsMax : nullptr, fam_zr);
}

// out_b[v] = sum over the subset's families whose sheet holds v's pixel of
// that family's partial of image b (family order). One wave per pixel, a
// lane per slice (64 slices at a time): the pixel's entries (position and
// slice range) are wave-uniform scalar loads, and a family whose elements
// miss a lane's slice at that pixel adds +0 without a load (its partial
// there, never written, is the empty sum).
constexpr int kRedBlock = 1024;
constexpr int kRedPx = kRedBlock / 64; // pixels per block, one per wave
constexpr size_t kRedTileMax = 65536;  // LDS bytes of the output tile
// With `tile` (LDS, [NB][nz][kRedPx] floats) the pixel's sums go to column
// px - px0 of the block's tile, which fam_reduce_store then writes out in
// runs of kRedPx consecutive pixels (instead of 4-byte stores one plane
// apart, which left most of each line to other blocks and moved up to 5x
// the bytes); without, straight to the images.
template <int NB>
__device__ __forceinline__ void fam_reduce_pixel(
  int px, const long long* __restrict__ off, const int2* __restrict__ ent,
  const float* __restrict__ part, int nxy, int nz, const VoxOut& out,
  float* tile = nullptr, int col = 0)
{
  const int lane = (int)threadIdx.x & 63;
  const long long a = off[px], b = off[px + 1];
  for (int z0 = 0; z0 < nz; z0 += 64)
  {
    const int z = z0 + lane;
    float acc[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) acc[i] = 0.0f;
    auto get = [&](int2 e) {
      Vec<NB> v;
      const unsigned r = (unsigned)e.y;
      if (z < nz && (unsigned)z >= (r & 0xffffu) && (unsigned)z < (r >> 16))
        v = Vec<NB>::load(part, (long long)e.x * nz + z);
      else
#pragma unroll
        for (int i = 0; i < NB; ++i) v.v[i] = 0.0f;
      return v;
    };
    long long k = a;
    for (; k + 4 <= b; k += 4) // four loads in flight, added in order
    {
      int2 e[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) e[u] = ent[k + u];
      Vec<NB> v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = get(e[u]);
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int i = 0; i < NB; ++i) acc[i] += v[u].v[i];
    }
    for (; k < b; ++k)
    {
      const Vec<NB> v = get(ent[k]);
#pragma unroll
      for (int i = 0; i < NB; ++i) acc[i] += v.v[i];
    }
    if (z < nz)
#pragma unroll
      for (int i = 0; i < NB; ++i)
      {
        if (tile)
          tile[(i * nz + z) * kRedPx + col] = acc[i];
        else if (out.p[i])
          out.p[i][(long long)z * nxy + px] = acc[i];
      }
  }
}

// The block's tile (pixels px0 .. px0 + kRedPx - 1) to the images.
template <int NB>
__device__ __forceinline__ void fam_reduce_store(const float* tile, int px0,
                                                 int nxy, int nz,
                                                 const VoxOut& out)
{
  const int n = NB * nz * kRedPx;
  for (int t = threadIdx.x; t < n; t += kRedBlock)
  {
    const int c = t % kRedPx, r = t / kRedPx;
    const int i = r / nz, z = r - i * nz;
    if (px0 + c < nxy && out.p[i])
      out.p[i][(long long)z * nxy + px0 + c] = tile[t];
  }
}

template <int NB>
__global__ __launch_bounds__(kRedBlock) void k_pc_fam_reduce(
  const long long* __restrict__ off, const int2* __restrict__ ent,
  const float* __restrict__ part, int nxy, int nz, VoxOut out, int stage)
{
  extern __shared__ float red_tile[];
  const int w = (int)threadIdx.x >> 6;
  const int px0 = (int)blockIdx.x * kRedPx;
  const int px = __builtin_amdgcn_readfirstlane(px0 + w);
  if (px < nxy)
    fam_reduce_pixel<NB>(px, off, ent, part, nxy, nz, out,
                         stage ? red_tile : nullptr, w);
  if (stage)
  {
    __syncthreads();
    fam_reduce_store<NB>(red_tile, px0, nxy, nz, out);
  }
}

// LDS bytes of the reduction's output tile for nb images of nz slices, or
// 0 when it does not fit (the sums then go straight to the images).
size_t red_tile_bytes(int nb, int nz)
{
  const size_t b = sizeof(float) * (size_t)nb * nz * kRedPx;
  return b <= kRedTileMax ? b : 0;
}

} // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct RowMatrix
{
  int device = 0;
  unsigned generation = 0;
  fir_vol_desc vol{};
  int walk_mode = 0;
  int pair_order = 0;
  int tang_major = 0;
  int lanes = 0;   // quad layout lanes (fir_set_cache_lanes)
  int zq = 1;      // z-slices interleaved in the gather layouts
  int fmt = 0;
  int partial = 0; // holds the rows of some subsets only (not shared)
  // Sheet form (tangential-major full matrices whose sheets fit in LDS).
  int n_fam = 0;                    // n_views * n_tang
  int n_pairs = 0;
  int* d_fam_row = nullptr;         // [n_fam + 1] first row of each family
  int* d_fam_off = nullptr;         // [n_fam + 1] first pixel of each sheet
  int* d_fam_xy = nullptr;          // sheet pixels (x + y nx)
  // per sheet pixel, the slices [lo, hi) of its entries that some element
  // crosses, lo | hi << 16 (entries outside are never read or written)
  unsigned* d_fam_zr = nullptr;
  unsigned* d_sel = nullptr;        // [n_slots] sheet index | length << 16
  // 1: sheet pixels ranked along each family's line, and every lane of the
  // quad layout meets its elements in rising pixel rank (checked at build)
  int line_order = 0;
  long long sheet_bytes = 0;
  long long n_sheet_px = 0;         // sheet pixels over all families
  // Per sheet entry (pixel, slice) of every family, the family's elements
  // through that voxel (row in family and length), whole quads stored as
  // ListQuads, in the slot order and round-interleaved groups of
  // k_fl_interleave.
  unsigned* d_fl_slot = nullptr;    // [n_sheet_px * nz] entry << 16 | quads
  unsigned* d_fl_gb = nullptr;      // [n_sheet_px * nz] group's first quad
  unsigned* d_fl = nullptr;         // [3 * fl_quads] ListQuads
  // per family, its slots holding a list or inside their pixel's slice
  // range (the slots before the ones never read)
  int* d_fl_nact = nullptr;
  long long fl_quads = 0;
  std::vector<unsigned> list_quads; // per family, from the sheet build
  int sheets_req = 0; // the context asked for the sheet form at build
  int n_views = 0;
  int n_rows = 0;
  long long n_elems = 0;            // path elements
  long long n_slots = 0;            // with each path padded to a multiple of 4
  int* d_view_row = nullptr;        // [n_views + 1]
  long long* d_view_elem = nullptr; // [n_views + 1]
  int* d_bin = nullptr;             // [n_rows]
  float* d_factor = nullptr;        // [n_rows]
  unsigned char* d_yrow = nullptr;  // [n_rows] 1: indices in the y-fastest layout
  long long* d_off = nullptr;       // [n_rows + 1]
  int* d_idx = nullptr;             // [n_slots]
  void* d_len = nullptr;            // [n_slots]
  std::vector<int> view_row;
  std::vector<long long> view_elem;

  long long bytes() const
  {
    return (long long)(n_views + 1) * (4 + 8) + (long long)n_rows * (4 + 4 + 1 + 8) +
      8 + n_slots * (4 + (long long)len_bytes(fmt)) + sheet_bytes +
      (d_fl ? 8 * n_sheet_px * vol.nz + 12 * fl_quads + 4LL * n_fam : 0);
  }

  ~RowMatrix()
  {
    (void)hipSetDevice(device);
    (void)hipFree(d_view_row);
    (void)hipFree(d_view_elem);
    (void)hipFree(d_bin);
    (void)hipFree(d_factor);
    (void)hipFree(d_yrow);
    (void)hipFree(d_off);
    (void)hipFree(d_idx);
    (void)hipFree(d_len);
    (void)hipFree(d_fam_row);
    (void)hipFree(d_fam_off);
    (void)hipFree(d_fam_xy);
    (void)hipFree(d_fam_zr);
    (void)hipFree(d_sel);
    (void)hipFree(d_fl_slot);
    (void)hipFree(d_fl_gb);
    (void)hipFree(d_fl);
    (void)hipFree(d_fl_nact);
  }
};

struct fir_cache
{
  fir_ctx* ctx = nullptr;
  std::shared_ptr<RowMatrix> rm;
  fir_vol_desc vol{};
  long long nvox = 0;
  int n_subsets = 0;
  int n_cached = 0;            // subsets 0 .. n_cached-1 are held
  long long* d_voff = nullptr; // [n_cached][nvox + 1], global element offsets
  int* d_trow = nullptr;       // [t_slots]
  void* d_tlen = nullptr;      // [t_slots]
  long long t_slots = 0;
  float* d_w = nullptr;        // per-row weights, [n_rows * 8]
  size_t w_floats = 0;
  float* d_x = nullptr;        // interleaved images, [nvox * 8]
  size_t x_floats = 0;
  float* d_xy = nullptr;       // the same in the y-fastest layout
  int fwd_lanes = 8;
  int bp_lanes = 16;
  int xcd_map = 0;
  // Backprojections of one or two images walk the voxels z fastest (with
  // more images the scattered per-image stores cost more than it saves).
  int zfast = 1;
  // One-image OSEM line integrals from the matrix's sheet form, if built.
  int sheets = 1;
  // Family backprojection (with the sheet form): per subset and pixel the
  // sheet entries holding it, and the partial sheets of one launch.
  long long* d_fr_off = nullptr; // [n_subsets * (nxy + 1) + 1]
  // [sheet entries] positions in the subset and slice ranges (lo | hi << 16)
  int2* d_fr_ent = nullptr;
  int* d_fam_loc = nullptr;      // [n_fam] first position of each family
  // Per subset, the families the sheet kernels run over (those with rows
  // or sheet pixels), most work first; subset s's at ford_first[s].
  int* d_fam_ord = nullptr;
  std::vector<int> ford_first;
  std::vector<int> h_ford;        // the same order on the host
  std::vector<long long> h_fcost; // each family's work (the sort key)
  // partial sheets of one launch, up to 8 images: [subset entries * nz * 8]
  float* d_part = nullptr;
  long long part_floats = 0;
  long long fr_bytes = 0;
  int family_bp = 1;
  // Batches of 4 or 8 images take their line integrals from slab-staged
  // sheets (when the matrix's sheet pixels are in line order).
  int slabs = 1;

  ~fir_cache()
  {
    (void)hipSetDevice(rm ? rm->device : 0);
    (void)hipFree(d_voff);
    (void)hipFree(d_trow);
    (void)hipFree(d_tlen);
    (void)hipFree(d_w);
    (void)hipFree(d_x);
    (void)hipFree(d_xy);
    (void)hipFree(d_fr_off);
    (void)hipFree(d_fr_ent);
    (void)hipFree(d_fam_loc);
    (void)hipFree(d_fam_ord);
    (void)hipFree(d_part);
  }
};

namespace
{

// The gather layouts of a cache's images.
XY cache_xy(const fir_cache* cache)
{
  return XY{cache->vol.nx, cache->vol.ny, cache->rm->zq};
}

// Device buffers of one build step, freed on every exit path.
struct DevBufs
{
  std::vector<void*> p;
  ~DevBufs()
  {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  int get(T** out, size_t count, const char* what)
  {
    *out = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc((void**)out, count * sizeof(T)) != hipSuccess)
    {
      (void)hipGetLastError();
      return fail(FIR_E_NOMEM, "path cache: allocation of %zu bytes for %s failed",
                  count * sizeof(T), what);
    }
    p.push_back(*out);
    return FIR_OK;
  }
  void release(void* q)
  {
    for (size_t i = 0; i < p.size(); ++i)
      if (p[i] == q)
      {
        (void)hipFree(q);
        p.erase(p.begin() + (long)i);
        return;
      }
  }
};

// Device buffers reused by the steps of one transpose: a step asks for at
// least `count` elements of buffer `slot` and gets that slot's buffer,
// reallocated only when it is too small. This avoids a multi-GB
// hipMalloc/hipFree pair per subset: freed device memory is cleared before
// it is handed out again, which costs seconds once most of it is in use.
enum PoolSlot
{
  kPoolRowCounts,
  kPoolRowPos,
  kPoolRowScan,
  kPoolKeys0,
  kPoolKeys1,
  kPoolVals0,
  kPoolVals1,
  kPoolSort,
  kPoolPrefix,
  kPoolVoxCounts,
  kPoolVoxOffsets,
  kPoolVoxScan,
  kPoolSlots
};

struct Pool
{
  void* p[kPoolSlots] = {};
  size_t bytes[kPoolSlots] = {};
  ~Pool()
  {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  int get(T** out, size_t count, PoolSlot slot, const char* what)
  {
    const size_t need = (count ? count : 1) * sizeof(T);
    if (bytes[slot] < need)
    {
      (void)hipFree(p[slot]);
      p[slot] = nullptr;
      bytes[slot] = 0;
      if (hipMalloc(&p[slot], need) != hipSuccess)
      {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(FIR_E_NOMEM,
                    "path cache: allocation of %zu bytes for %s failed", need,
                    what);
      }
      bytes[slot] = need;
    }
    *out = static_cast<T*>(p[slot]);
    return FIR_OK;
  }
};

template <class T>
int dev_alloc(T** out, size_t count, const char* what)
{
  *out = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc((void**)out, count * sizeof(T)) != hipSuccess)
  {
    (void)hipGetLastError();
    return fail(FIR_E_NOMEM, "path cache: allocation of %zu bytes for %s failed",
                count * sizeof(T), what);
  }
  return FIR_OK;
}

bool same_vol(const fir_vol_desc& a, const fir_vol_desc& b)
{
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dx == b.dx &&
    a.dy == b.dy && a.dz == b.dz && a.ox == b.ox && a.oy == b.oy &&
    a.oz == b.oz;
}

// LOR decode of the cache's rows (see cache_lor): the context's pair
// order and tangential-major flag (fir_set_lor_order).
LorGeom lor_geom_views(const fir_ctx* c)
{
  LorGeom lg = lor_geom(c);
  lg.row_mode = fir::ROWS_VIEW_MAJOR;
  return lg;
}

template <int F>
int fill_rows(fir_ctx* c, const Grid& g, const LorGeom& lg, long long n,
              long long per_view, const int* count, const float* maxlen, const long long* eoff,
              RowMatrix* rm)
{
  const int lanes = c->cache_lanes[0];
  const int stage = lanes <= 8 ? 1 : 0;
  const size_t lds = stage ? (size_t)kBlock * 4 * lanes *
      (sizeof(int) + sizeof(typename Len<F>::T)) : 0;
  hipLaunchKernelGGL(k_pc_fill<F>, dim3(blocks_for(n)), dim3(kBlock), lds,
                     c->stream, g, lg, n, per_view, count, maxlen, eoff,
                     XY{g.row, g.slice / g.row, rm->zq}, lanes, stage,
                     rm->d_idx, (typename Len<F>::T*)rm->d_len);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

// A partial cache: subsets 0 .. n_cached-1 of S, n_cached at most
// max_cached (<= 0: no cap), the largest count whose device memory (build
// temporaries included) fits in max_bytes (<= 0: no limit) and in the free
// device memory.
struct PartialPlan
{
  int S = 0;
  int max_cached = 0;
  long long max_bytes = 0;
  long long nvox = 0;
  long long image_bytes = 0; // the cache's image copies
  int n_cached = 0;          // chosen
};

// Device bytes of subsets 0 .. k-1 of a cache, from per-subset totals.
int plan_subsets(fir_ctx* c, const int* count, long long n, long long per_view,
                 int fmt, PartialPlan* plan)
{
  const int V = c->n_views, S = plan->S;
  DevBufs tmp;
  unsigned long long* d_tot;
  int rc;
  if ((rc = tmp.get(&d_tot, 3 * (size_t)V, "view totals"))) return rc;
  hipLaunchKernelGGL(k_pc_view_totals, dim3(V), dim3(kBlock), 0, c->stream,
                     count, per_view, d_tot);
  FIR_CHECK_LAUNCH();
  std::vector<unsigned long long> tot(3 * (size_t)V);
  FIR_HIP_CHECK(hipMemcpyAsync(tot.data(), d_tot, tot.size() * sizeof(tot[0]),
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  std::vector<long long> slots(S, 0), elems(S, 0), rows(S, 0);
  for (int v = 0; v < V; ++v)
  {
    slots[v % S] += (long long)tot[3 * (size_t)v];
    elems[v % S] += (long long)tot[3 * (size_t)v + 1];
    rows[v % S] += (long long)tot[3 * (size_t)v + 2];
  }
  size_t free_b = 0, total_b = 0;
  FIR_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  const long long lb = (long long)len_bytes(fmt);
  const long long held = 20LL * n; // count, maxlen, row ids, offsets
  const long long avail = (long long)free_b - (1LL << 30);
  const int cap = plan->max_cached > 0 ? std::min(plan->max_cached, S) : S;
  long long R = 0, SL = 0, E = 0, maxSL = 0, cost1 = 0;
  int k = 0;
  for (int s = 0; s < cap; ++s)
  {
    R += rows[s];
    SL += slots[s];
    E += elems[s];
    maxSL = std::max(maxSL, slots[s]);
    const long long cost =
      R * (4 + 4 + 1 + 8) + SL * (4 + lb) + (V + 1LL) * 12 +         // rows
      (E + 3LL * (s + 1) * plan->nvox) * (4 + lb) +                     // lists
      (s + 1LL) * (plan->nvox + 1) * 8 +                                // offsets
      R * kMaxCacheBatch * 4 + plan->image_bytes +                      // weights
      std::max(held, 24 * std::min(c->cache_chunk, maxSL) + 16 * plan->nvox);
    if (s == 0) cost1 = cost;
    if ((plan->max_bytes > 0 && cost > plan->max_bytes) || cost - held > avail)
      break;
    k = s + 1;
  }
  if (k == 0)
    return fail(FIR_E_NOMEM,
                "path cache: subset 0 alone needs %.2f GB of device memory "
                "(limit %.2f GB, %.2f GB free)",
                cost1 / 1e9, plan->max_bytes > 0 ? plan->max_bytes / 1e9 : 0.0,
                free_b / 1e9);
  plan->n_cached = k;
  return FIR_OK;
}

// The sheet form of a filled matrix whose family first rows are set
// (d_fam_row). Sheets that do not fit in LDS, or a device too full for the
// 2-byte indices, leave the matrix without it (the projectors then gather
// from the image); only HIP failures are errors.
int build_sheets(fir_ctx* c, RowMatrix* rm)
{
  const fir_vol_desc& v = rm->vol;
  const int nxy = v.nx * v.ny, nz = v.nz;
  const int words = (nxy + 63) / 64;
  const int F = rm->n_fam;
  // 16-bit lengths only: an element is a 16-bit sheet index and the
  // length's 16 bits in one word.
  if ((long long)(v.nx + v.ny) * nz > 65535 || rm->n_rows == 0 ||
      len_bytes(rm->fmt) != 2)
    return FIR_OK;
  size_t free_b = 0, total_b = 0;
  FIR_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  const long long temp = (long long)F * words * 8 + 8LL * F;
  if ((long long)free_b < temp + rm->n_slots * 4 + (2LL << 30)) return FIR_OK;
  DevBufs tmp;
  unsigned long long* bits;
  int* n_px;
  int rc;
  if ((rc = tmp.get(&bits, (size_t)F * words, "sheet bitmaps")) ||
      (rc = tmp.get(&n_px, (size_t)F, "sheet sizes")))
    return rc;
  const XY xy{v.nx, v.ny, rm->zq};
  const int lb = (int)len_bytes(rm->fmt);
  const size_t mark_lds = sizeof(unsigned long long) * words;
  if (mark_lds > 65536) return FIR_OK;
  hipLaunchKernelGGL(k_sh_mark_fam, dim3(F), dim3(kBlock), mark_lds,
                     c->stream, rm->d_fam_row, rm->d_off, rm->d_idx,
                     rm->d_len, lb, rm->d_yrow, xy, words, bits, n_px);
  FIR_CHECK_LAUNCH();
  std::vector<int> npx(F), off(F + 1, 0);
  FIR_HIP_CHECK(hipMemcpyAsync(npx.data(), n_px, sizeof(int) * F,
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  int max_px = 0;
  for (int f = 0; f < F; ++f)
  {
    off[f + 1] = off[f] + npx[f];
    max_px = std::max(max_px, npx[f]);
  }
  const size_t enc_lds = 8 * (size_t)words +
    4 * (size_t)((words + 1) & ~1) + 12 * (size_t)max_px +
    2 * (size_t)((max_px + 1) & ~1) + 2 * (size_t)(max_px * nz + 1);
  if ((long long)max_px * nz > kSheetMaxFloats || enc_lds > 65536)
    return FIR_OK;
  if ((rc = dev_alloc(&rm->d_fam_off, (size_t)F + 1, "sheet offsets")) ||
      (rc = dev_alloc(&rm->d_fam_xy, (size_t)off[F], "sheet pixels")) ||
      (rc = dev_alloc(&rm->d_fam_zr, (size_t)std::max(off[F], 1),
                      "sheet slice ranges")) ||
      (rc = dev_alloc(&rm->d_sel, (size_t)rm->n_slots, "sheet elements")))
  {
    (void)hipFree(rm->d_fam_off);
    (void)hipFree(rm->d_fam_xy);
    (void)hipFree(rm->d_fam_zr);
    rm->d_fam_off = nullptr;
    rm->d_fam_xy = nullptr;
    rm->d_fam_zr = nullptr;
    rm->d_sel = nullptr;
    return rc == FIR_E_NOMEM ? FIR_OK : rc;
  }
  FIR_HIP_CHECK(hipMemcpyAsync(rm->d_fam_off, off.data(), sizeof(int) * (F + 1),
                               hipMemcpyHostToDevice, c->stream));
  // Pixels ranked along each family's line (the batches' slab staging
  // needs it; one-image passes give the same sums in either order).
  Grid g;
  if ((rc = make_grid(c, &v, &g))) return rc;
  const int n_tang = F / rm->n_views;
  const bool line = F % rm->n_views == 0;
  unsigned* list_quads;
  if ((rc = tmp.get(&list_quads, (size_t)F, "family list sizes"))) return rc;
  FIR_HIP_CHECK(hipMemsetAsync(list_quads, 0, sizeof(unsigned) * F,
                               c->stream));
  hipLaunchKernelGGL(k_sh_encode_fam, dim3(F), dim3(kBlock), enc_lds,
                     c->stream, lor_geom_views(c), g, n_tang, rm->d_fam_row,
                     rm->d_off, rm->d_idx, rm->d_len, lb, rm->d_yrow, xy, nz,
                     words, bits, rm->d_fam_off, line ? 1 : 0, rm->d_fam_xy,
                     rm->d_fam_zr, rm->d_sel, list_quads);
  FIR_CHECK_LAUNCH();
  rm->list_quads.resize(F);
  FIR_HIP_CHECK(hipMemcpyAsync(rm->list_quads.data(), list_quads,
                               sizeof(unsigned) * F, hipMemcpyDeviceToHost,
                               c->stream));
  rm->line_order = 0;
  if (line)
  {
    int* ok;
    if ((rc = tmp.get(&ok, 1, "sheet order flag"))) return rc;
    const int one = 1;
    FIR_HIP_CHECK(hipMemcpyAsync(ok, &one, sizeof(int), hipMemcpyHostToDevice,
                                 c->stream));
    hipLaunchKernelGGL(k_sh_check_order,
                       dim3(blocks_for((long long)rm->n_rows * rm->lanes)),
                       dim3(kBlock), 0, c->stream, rm->n_rows, rm->d_off,
                       rm->d_sel, rm->lanes, nz, ok);
    FIR_CHECK_LAUNCH();
    FIR_HIP_CHECK(hipMemcpyAsync(&rm->line_order, ok, sizeof(int),
                                 hipMemcpyDeviceToHost, c->stream));
  }
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  rm->sheet_bytes = 4LL * (F + 1) + 8LL * off[F] + 4LL * rm->n_slots;
  rm->n_sheet_px = off[F];
  return FIR_OK;
}

// The per-entry element lists of every family (the family backprojection's
// input), for a matrix with the sheet form. Families with more rows than
// the fused kernel holds, lists beyond 2^32 quads, row steps within a quad
// beyond ListQuad's 7 bits, or a device too full leave the matrix without
// them; only HIP failures are errors.
int build_family_lists(fir_ctx* c, RowMatrix* rm)
{
  const int F = rm->n_fam, nz = rm->vol.nz;
  if (!rm->d_sel || rm->n_pairs > kFamRowsMax || rm->n_sheet_px <= 0)
    return FIR_OK;
  DevBufs tmp;
  long long* fbase;
  int* flags; // [0]: a round table overflowed, [1]: a quad does not pack
  int rc;
  if ((rc = tmp.get(&fbase, (size_t)F, "family list bases")) ||
      (rc = tmp.get(&flags, 2, "family list flags")))
    return rc;
  // the quads of every family's lists, counted by the sheet build
  const std::vector<unsigned>& tot = rm->list_quads;
  if ((int)tot.size() != F) return FIR_OK;
  std::vector<long long> base(F);
  long long quads = 0;
  for (int f = 0; f < F; ++f)
  {
    base[f] = quads;
    quads += tot[f];
  }
  const long long entries = rm->n_sheet_px * nz;
  if (quads >= (1LL << 32)) return FIR_OK;
  const long long stage_q = std::min(quads, kFlStageQuads);
  // runs of families whose quads fit the staging buffer (a family holds far
  // fewer: n_st * 512)
  std::vector<int> runs{0};
  for (int f0 = 0; f0 < F;)
  {
    int f1 = f0 + 1;
    while (f1 < F && base[f1] + tot[f1] - base[f0] <= stage_q) ++f1;
    if ((f1 < F ? base[f1] : quads) - base[f0] > stage_q)
      return fail(FIR_E_NOMEM, "family lists: a family exceeds the staging");
    runs.push_back(f0 = f1);
  }
  auto drop = [&]() {
    (void)hipFree(rm->d_fl_slot);
    (void)hipFree(rm->d_fl_gb);
    (void)hipFree(rm->d_fl_nact);
    (void)hipFree(rm->d_fl);
    rm->d_fl_slot = nullptr;
    rm->d_fl_gb = nullptr;
    rm->d_fl_nact = nullptr;
    rm->d_fl = nullptr;
  };
  auto alloc_lists = [&]() {
    int r;
    if ((r = dev_alloc(&rm->d_fl_slot, (size_t)entries,
                       "family list slots")) ||
        (r = dev_alloc(&rm->d_fl_gb, (size_t)entries, "family list groups")) ||
        (r = dev_alloc(&rm->d_fl_nact, (size_t)F, "family list actives")) ||
        (r = dev_alloc(&rm->d_fl, (size_t)(3 * std::max(quads, 1LL)),
                       "family lists")))
      drop();
    return r;
  };
  // a run's quads from the staging buffer into their packed form
  auto pack = [&](unsigned* stage, int i) {
    const long long q0 = base[runs[i]];
    const long long nq = (runs[i + 1] < F ? base[runs[i + 1]] : quads) - q0;
    if (nq > 0)
      hipLaunchKernelGGL(k_fl_pack, dim3((unsigned)((nq + kBlock - 1) / kBlock)),
                         dim3(kBlock), 0, c->stream, stage, nq,
                         rm->d_fl + 3 * q0, flags + 1);
    return q0;
  };
  auto finish = [&](int* ovf) -> int {
    FIR_HIP_CHECK(hipMemcpyAsync(ovf, flags, 2 * sizeof(int),
                                 hipMemcpyDeviceToHost, c->stream));
    FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
    return FIR_OK;
  };
  size_t free_b = 0, total_b = 0;
  FIR_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  unsigned* stage;
  if ((long long)free_b >= 12 * quads + 16 * stage_q + 8 * entries + (3LL << 30))
  {
    // one pass per family into the final layout, run by run through the
    // staging buffer
    if ((rc = tmp.get(&stage, (size_t)(4 * std::max(stage_q, 1LL)),
                      "family list staging")))
      return rc == FIR_E_NOMEM ? FIR_OK : rc;
    if ((rc = alloc_lists())) return rc == FIR_E_NOMEM ? FIR_OK : rc;
    FIR_HIP_CHECK(hipMemcpyAsync(fbase, base.data(), 8LL * F,
                                 hipMemcpyHostToDevice, c->stream));
    FIR_HIP_CHECK(hipMemsetAsync(flags, 0, 2 * sizeof(int), c->stream));
    FIR_HIP_CHECK(hipMemsetAsync(rm->d_fl_nact, 0, 4LL * F, c->stream));
    for (size_t i = 0; i + 1 < runs.size(); ++i)
    {
      const long long q0 = base[runs[i]];
      hipLaunchKernelGGL(k_fl_build, dim3(runs[i + 1] - runs[i]),
                         dim3(kFlBlock), 0, c->stream, runs[i],
                         rm->d_fam_row, rm->d_fam_off, nz, rm->d_fam_zr,
                         rm->d_off, rm->d_sel, fbase, rm->d_fl_slot,
                         rm->d_fl_gb, stage, q0, rm->d_fl_nact, flags);
      FIR_CHECK_LAUNCH();
      pack(stage, (int)i);
      FIR_CHECK_LAUNCH();
    }
    int ovf[2] = {0, 0};
    if ((rc = finish(ovf))) return rc;
    if (!ovf[0] && !ovf[1])
    {
      rm->fl_quads = quads;
      return FIR_OK;
    }
    drop();
    if (!ovf[0]) return FIR_OK; // row steps beyond 7 bits
    // some family's round table is too large for LDS: the two-pass route
    tmp.release(stage);
    FIR_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  }
  // the contiguous sorted lists go into a full-size buffer first; runs of
  // families are then interleaved through the staging buffer and packed
  if ((long long)free_b <
      16 * (quads + stage_q) + 12 * quads + 12 * entries + (3LL << 30))
    return FIR_OK;
  unsigned* cl_off;
  unsigned* cl;
  if ((rc = tmp.get(&cl_off, (size_t)entries + 1, "family list offsets")) ||
      (rc = tmp.get(&cl, (size_t)(4 * std::max(quads, 1LL)),
                    "family contiguous lists")) ||
      (rc = tmp.get(&stage, (size_t)(4 * std::max(stage_q, 1LL)),
                    "family list staging")))
    return rc == FIR_E_NOMEM ? FIR_OK : rc;
  if ((rc = alloc_lists())) return rc == FIR_E_NOMEM ? FIR_OK : rc;
  FIR_HIP_CHECK(hipMemsetAsync(flags, 0, 2 * sizeof(int), c->stream));
  FIR_HIP_CHECK(hipMemsetAsync(rm->d_fl_nact, 0, 4LL * F, c->stream));
  const unsigned end = (unsigned)quads;
  FIR_HIP_CHECK(hipMemcpyAsync(fbase, base.data(), 8LL * F,
                               hipMemcpyHostToDevice, c->stream));
  FIR_HIP_CHECK(hipMemcpyAsync(cl_off + entries, &end, 4,
                               hipMemcpyHostToDevice, c->stream));
  FIR_HIP_CHECK(hipMemsetAsync(cl, 0, 16 * quads, c->stream));
  hipLaunchKernelGGL(k_fl_fill, dim3(F), dim3(kFlBlock), 0, c->stream,
                     rm->d_fam_row, rm->d_fam_off, nz, rm->d_off, rm->d_sel,
                     fbase, cl_off, cl);
  FIR_CHECK_LAUNCH();
  for (size_t i = 0; i + 1 < runs.size(); ++i)
  {
    const long long q0 = base[runs[i]];
    hipLaunchKernelGGL(k_fl_interleave, dim3(runs[i + 1] - runs[i]),
                       dim3(kFlBlock), 0, c->stream, runs[i], rm->d_fam_off,
                       nz, rm->d_fam_zr, fbase, cl_off, cl, rm->d_fl_slot,
                       rm->d_fl_gb, stage, rm->d_fl_nact, q0);
    FIR_CHECK_LAUNCH();
    pack(stage, (int)i);
    FIR_CHECK_LAUNCH();
  }
  int ovf[2] = {0, 0};
  if ((rc = finish(ovf))) return rc;
  if (ovf[1])
  {
    drop();
    return FIR_OK;
  }
  rm->fl_quads = quads;
  return FIR_OK;
}

// Path slots of every family's rows.
__global__ void k_fam_work(int n_fam, const int* __restrict__ fam_row,
                           const long long* __restrict__ off,
                           long long* __restrict__ work)
{
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n_fam) work[f] = off[fam_row[f + 1]] - off[fam_row[f]];
}

// The order in which the sheet kernels (one block per family) take each
// subset's families (views k * S + s): only families with rows or sheet
// pixels, the most work (path slots plus sheet entries) first, so that a
// launch ends on its smallest families instead of on whichever large ones
// come last in view order, and no block is spent on an empty family.
int build_family_order(fir_ctx* c, fir_cache* cache)
{
  const RowMatrix* rm = cache->rm.get();
  if (!rm->d_sel || cache->d_fam_ord) return FIR_OK;
  const int F = rm->n_fam, S = cache->n_subsets, nz = cache->vol.nz;
  const int n_tang = F / rm->n_views;
  DevBufs tmp;
  long long* work;
  int rc;
  if ((rc = tmp.get(&work, (size_t)F, "family work"))) return rc;
  hipLaunchKernelGGL(k_fam_work, dim3((F + kBlock - 1) / kBlock),
                     dim3(kBlock), 0, c->stream, F, rm->d_fam_row, rm->d_off,
                     work);
  FIR_CHECK_LAUNCH();
  std::vector<long long> hw(F);
  std::vector<int> frow(F + 1), foff(F + 1);
  FIR_HIP_CHECK(hipMemcpyAsync(hw.data(), work, 8LL * F,
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipMemcpyAsync(frow.data(), rm->d_fam_row, 4LL * (F + 1),
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipMemcpyAsync(foff.data(), rm->d_fam_off, 4LL * (F + 1),
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  std::vector<int> ord, first(S + 1);
  ord.reserve(F);
  for (int s0 = 0; s0 < S; ++s0)
  {
    first[s0] = (int)ord.size();
    const int vpr = fir::views_in_subset(rm->n_views, S, s0);
    for (int k = 0; k < vpr; ++k)
      for (int t = 0; t < n_tang; ++t)
      {
        const int f = (k * S + s0) * n_tang + t;
        if (frow[f + 1] > frow[f] || foff[f + 1] > foff[f]) ord.push_back(f);
      }
    auto cost = [&](int f) {
      return hw[f] + (long long)(foff[f + 1] - foff[f]) * nz;
    };
    std::stable_sort(ord.begin() + first[s0], ord.end(),
                     [&](int a, int b) { return cost(a) > cost(b); });
  }
  first[S] = (int)ord.size();
  if ((rc = dev_alloc(&cache->d_fam_ord, std::max(ord.size(), (size_t)1),
                      "family order")))
    return rc;
  if (!ord.empty())
    FIR_HIP_CHECK(hipMemcpyAsync(cache->d_fam_ord, ord.data(),
                                 4 * ord.size(), hipMemcpyHostToDevice,
                                 c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  cache->ford_first = first;
  cache->h_ford = ord;
  cache->h_fcost.resize(F);
  for (int f = 0; f < F; ++f)
    cache->h_fcost[f] = hw[f] + (long long)(foff[f + 1] - foff[f]) * nz;
  return FIR_OK;
}

// Tables of the family backprojection (k_pc_osem_fam + k_pc_fam_reduce) for
// a cache over a matrix with family lists: per subset and pixel, the sheet
// entries that hold it, in family order, and the partial sheets. A matrix
// without the lists, or a device too full, leaves the cache on the
// transposed-list backprojection; only HIP failures are errors.
int build_family_bp(fir_ctx* c, fir_cache* cache)
{
  const RowMatrix* rm = cache->rm.get();
  if (!rm->d_fl) return FIR_OK;
  const int S = cache->n_subsets;
  const int nxy = cache->vol.nx * cache->vol.ny, nz = cache->vol.nz;
  const long long total = rm->n_sheet_px;
  const long long n_keys = (long long)S * (nxy + 1);
  const int n_tang = rm->n_fam / rm->n_views;
  // each family's first position among its subset's sheet pixels (views
  // v = k * S + s in order, tangential bins in order), and the largest
  // subset: the partial sheets of a launch hold one subset
  std::vector<int> foff(rm->n_fam + 1), floc(rm->n_fam);
  FIR_HIP_CHECK(hipMemcpyAsync(foff.data(), rm->d_fam_off,
                               4LL * (rm->n_fam + 1), hipMemcpyDeviceToHost,
                               c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  long long max_loc = 0;
  for (int s0 = 0; s0 < S; ++s0)
  {
    long long loc = 0;
    for (int view = s0; view < rm->n_views; view += S)
      for (int t = 0; t < n_tang; ++t)
      {
        const int f = view * n_tang + t;
        floc[f] = (int)loc;
        loc += foff[f + 1] - foff[f];
      }
    max_loc = std::max(max_loc, loc);
  }
  const long long part_floats = max_loc * nz * kMaxCacheBatch;
  const long long bytes =
    8 * (n_keys + 1) + 8 * total + 4LL * rm->n_fam + 4 * part_floats;
  size_t free_b = 0, total_b = 0;
  FIR_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  if ((long long)free_b < bytes + 4 * n_keys + (2LL << 30)) return FIR_OK;
  DevBufs tmp;
  unsigned* cnt;
  int rc;
  if ((rc = tmp.get(&cnt, (size_t)n_keys, "family table counts"))) return rc;
  if ((rc = dev_alloc(&cache->d_fr_off, (size_t)n_keys + 1,
                      "family table offsets")) ||
      (rc = dev_alloc(&cache->d_fr_ent, (size_t)std::max(total, 1LL),
                      "family table")) ||
      (rc = dev_alloc(&cache->d_fam_loc, (size_t)rm->n_fam,
                      "family positions")) ||
      (rc = dev_alloc(&cache->d_part, (size_t)std::max(part_floats, 1LL),
                      "family partial sheets")))
  {
    (void)hipFree(cache->d_fr_off);
    (void)hipFree(cache->d_fr_ent);
    (void)hipFree(cache->d_fam_loc);
    cache->d_fr_off = nullptr;
    cache->d_fr_ent = nullptr;
    cache->d_fam_loc = nullptr;
    cache->d_part = nullptr;
    return rc == FIR_E_NOMEM ? FIR_OK : rc;
  }
  FIR_HIP_CHECK(hipMemcpyAsync(cache->d_fam_loc, floc.data(),
                               4LL * rm->n_fam, hipMemcpyHostToDevice,
                               c->stream));
  FIR_HIP_CHECK(hipMemsetAsync(cnt, 0, 4 * n_keys, c->stream));
  hipLaunchKernelGGL(k_fr_count, dim3(blocks_for(total)), dim3(kBlock), 0,
                     c->stream, rm->d_fam_off, rm->d_fam_xy, rm->n_fam,
                     (int)total, n_tang, S, nxy, cnt);
  FIR_CHECK_LAUNCH();
  std::vector<unsigned> h(n_keys);
  std::vector<long long> off(n_keys + 1, 0);
  FIR_HIP_CHECK(hipMemcpyAsync(h.data(), cnt, 4 * n_keys,
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  for (long long k = 0; k < n_keys; ++k) off[k + 1] = off[k] + h[k];
  FIR_HIP_CHECK(hipMemcpyAsync(cache->d_fr_off, off.data(), 8 * (n_keys + 1),
                               hipMemcpyHostToDevice, c->stream));
  FIR_HIP_CHECK(hipMemsetAsync(cnt, 0, 4 * n_keys, c->stream));
  hipLaunchKernelGGL(k_fr_fill, dim3(blocks_for(total)), dim3(kBlock), 0,
                     c->stream, rm->d_fam_off, rm->d_fam_xy, rm->d_fam_zr,
                     cache->d_fam_loc, rm->n_fam, (int)total, n_tang, S, nxy,
                     cache->d_fr_off, cnt, cache->d_fr_ent);
  FIR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_fr_sort, dim3(blocks_for(n_keys)), dim3(kBlock), 0,
                     c->stream, cache->d_fr_off, n_keys, cache->d_fr_ent);
  FIR_CHECK_LAUNCH();
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  cache->fr_bytes = bytes;
  cache->part_floats = part_floats;
  return FIR_OK;
}

int build_rows(fir_ctx* c, const fir_vol_desc* v, int fmt,
               std::shared_ptr<RowMatrix>* out, PartialPlan* plan = nullptr)
{
  Grid g;
  int rc = make_grid(c, v, &g);
  if (rc) return rc;
  const long long n = c->n_bins;
  if ((rc = check_work(n))) return rc;
  const LorGeom lg = lor_geom_views(c);
  const long long per_view = (long long)c->n_pairs * c->n_tang;
  if (per_view * c->n_views != n)
    return fail(FIR_E_STATE, "path cache: inconsistent sinogram geometry");

  DevBufs tmp;
  int* count;
  float* maxlen;
  int* rowid;
  long long* eoff;
  if ((rc = tmp.get(&count, n, "element counts"))) return rc;
  if ((rc = tmp.get(&maxlen, n, "longest elements"))) return rc;
  if ((rc = tmp.get(&rowid, n, "row ids"))) return rc;
  if ((rc = tmp.get(&eoff, n, "element offsets"))) return rc;

  hipLaunchKernelGGL(k_pc_stats, dim3(blocks_for(n)), dim3(kBlock), 0,
                     c->stream, g, lg, n, per_view, count, maxlen);
  FIR_CHECK_LAUNCH();
  if (plan)
  {
    if ((rc = plan_subsets(c, count, n, per_view, fmt, plan))) return rc;
    if (plan->n_cached < plan->S)
    {
      hipLaunchKernelGGL(k_pc_keep_subsets, dim3(blocks_for(n)), dim3(kBlock),
                         0, c->stream, count, n, per_view, plan->S,
                         plan->n_cached);
      FIR_CHECK_LAUNCH();
    }
  }

  // Exclusive scans: row id of each crossing LOR, first element of each LOR.
  hipcub::TransformInputIterator<int, HitFlag, const int*> hits(count,
                                                               HitFlag());
  hipcub::TransformInputIterator<long long, Pad4, const int*> counts64(count,
                                                                       Pad4());
  hipcub::TransformInputIterator<long long, ToI64, const int*> real64(count,
                                                                     ToI64());
  size_t tb1 = 0, tb2 = 0;
  FIR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb1, hits, rowid,
                                                 (int)n, c->stream));
  FIR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, counts64, eoff,
                                                 (int)n, c->stream));
  unsigned char* scan_tmp;
  if ((rc = tmp.get(&scan_tmp, std::max(tb1, tb2), "scan scratch"))) return rc;
  FIR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb1, hits, rowid,
                                                 (int)n, c->stream));
  FIR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb2, counts64, eoff,
                                                 (int)n, c->stream));

  auto rm = std::make_shared<RowMatrix>();
  rm->device = c->device;
  rm->generation = c->generation;
  rm->vol = *v;
  rm->walk_mode = c->walk_mode;
  rm->pair_order = c->lor_order;
  rm->tang_major = c->tang_major;
  rm->lanes = c->cache_lanes[0];
  rm->zq = c->cache_zq;
  rm->fmt = fmt;
  rm->partial = plan && plan->n_cached < plan->S ? 1 : 0;
  rm->sheets_req = c->cache_sheets;
  rm->n_views = c->n_views;
  if ((rc = dev_alloc(&rm->d_view_row, c->n_views + 1, "view rows"))) return rc;
  if ((rc = dev_alloc(&rm->d_view_elem, c->n_views + 1, "view elements")))
    return rc;

  // Totals first (they size the row and element arrays).
  long long* d_end;
  if ((rc = tmp.get(&d_end, 1, "totals"))) return rc;
  size_t tb3 = 0;
  FIR_HIP_CHECK(hipcub::DeviceReduce::Sum(nullptr, tb3, real64, d_end, (int)n,
                                          c->stream));
  unsigned char* red_tmp;
  if ((rc = tmp.get(&red_tmp, tb3, "reduction scratch"))) return rc;
  FIR_HIP_CHECK(hipcub::DeviceReduce::Sum(red_tmp, tb3, real64, d_end, (int)n,
                                          c->stream));
  long long real_total = 0;
  FIR_HIP_CHECK(hipMemcpyAsync(&real_total, d_end, sizeof(long long),
                               hipMemcpyDeviceToHost, c->stream));
  int last_row = 0, last_count = 0;
  long long last_eoff = 0;
  FIR_HIP_CHECK(hipMemcpyAsync(&last_row, rowid + n - 1, sizeof(int),
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipMemcpyAsync(&last_count, count + n - 1, sizeof(int),
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipMemcpyAsync(&last_eoff, eoff + n - 1, sizeof(long long),
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  rm->n_rows = last_row + (last_count > 0 ? 1 : 0);
  if (c->tang_major && !plan && c->cache_sheets)
  {
    // Families (view, tangential bin) are contiguous in this row order.
    rm->n_pairs = c->n_pairs;
    rm->n_fam = c->n_views * c->n_tang;
    if ((rc = dev_alloc(&rm->d_fam_row, (size_t)rm->n_fam + 1,
                        "family rows")))
      return rc;
    hipLaunchKernelGGL(k_sh_fam_rows, dim3(blocks_for(rm->n_fam + 1)),
                       dim3(kBlock), 0, c->stream, rowid, c->n_pairs,
                       rm->n_fam, rm->n_rows, rm->d_fam_row);
    FIR_CHECK_LAUNCH();
  }
  rm->n_elems = real_total;
  rm->n_slots = last_eoff + ((last_count + 3) & ~3);

  size_t free_b = 0, total_b = 0;
  FIR_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  const long long need = rm->bytes();
  if ((long long)free_b < need + (1LL << 30))
    return fail(FIR_E_NOMEM,
                "path cache: %lld elements need %.2f GB of device memory, "
                "%.2f GB free",
                rm->n_elems, need / 1e9, free_b / 1e9);
  if ((rc = dev_alloc(&rm->d_bin, rm->n_rows, "row bins"))) return rc;
  if ((rc = dev_alloc(&rm->d_factor, rm->n_rows, "row factors"))) return rc;
  if ((rc = dev_alloc(&rm->d_yrow, rm->n_rows, "row layouts"))) return rc;
  if ((rc = dev_alloc(&rm->d_off, (size_t)rm->n_rows + 1, "row offsets")))
    return rc;
  if ((rc = dev_alloc(&rm->d_idx, rm->n_slots, "element indices"))) return rc;
  if ((rc = dev_alloc((unsigned char**)&rm->d_len,
                      rm->n_slots * len_bytes(fmt), "element lengths")))
    return rc;
  // padding elements: index 0, length 0 (the fill writes them itself when
  // it stages whole rounds, lanes <= 8)
  if (c->cache_lanes[0] > 8)
  {
    FIR_HIP_CHECK(hipMemsetAsync(rm->d_idx, 0, sizeof(int) * rm->n_slots,
                                 c->stream));
    FIR_HIP_CHECK(hipMemsetAsync(rm->d_len, 0, rm->n_slots * len_bytes(fmt),
                                 c->stream));
  }

  hipLaunchKernelGGL(k_pc_views, dim3(blocks_for(c->n_views + 1)),
                     dim3(kBlock), 0, c->stream, count, rowid, eoff, n,
                     per_view, c->n_views, rm->d_view_row, rm->d_view_elem,
                     rm->d_off);
  FIR_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_pc_rows, dim3(blocks_for(n)), dim3(kBlock), 0,
                     c->stream, lg, n, per_view, count, maxlen, rowid, eoff,
                     fmt,
                     rm->d_bin, rm->d_factor, rm->d_off, rm->d_yrow);
  FIR_CHECK_LAUNCH();
  switch (fmt)
  {
  case FIR_LEN_SNORM16:
    rc = fill_rows<FIR_LEN_SNORM16>(c, g, lg, n, per_view, count, maxlen, eoff,
                                  rm.get());
    break;
  case FIR_LEN_FP16:
    rc = fill_rows<FIR_LEN_FP16>(c, g, lg, n, per_view, count, maxlen, eoff,
                                  rm.get());
    break;
  default:
    rc = fill_rows<FIR_LEN_FP32>(c, g, lg, n, per_view, count, maxlen, eoff,
                                  rm.get());
    break;
  }
  if (rc) return rc;
  rm->view_row.resize(c->n_views + 1);
  rm->view_elem.resize(c->n_views + 1);
  FIR_HIP_CHECK(hipMemcpyAsync(rm->view_row.data(), rm->d_view_row,
                               sizeof(int) * (c->n_views + 1),
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipMemcpyAsync(rm->view_elem.data(), rm->d_view_elem,
                               sizeof(long long) * (c->n_views + 1),
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (rm->d_fam_row && ((rc = build_sheets(c, rm.get())) ||
                        (rc = build_family_lists(c, rm.get()))))
    return rc;
  *out = rm;
  return FIR_OK;
}

template <int F>
int transpose_chunk(fir_ctx* c, const RowMatrix* rm, fir_cache* cache,
                    const SubsetRows& sr, const int* d_prefix, int vpr,
                    int max_rows, int n_rows_s, int v0, int v1,
                    const long long* voff_p, const long long* voff_u,
                    long long n_ch,
                    const std::function<int(const unsigned*, long long)>&
                      offsets_from_sorted,
                    Pool& tmp)
{
  // n_ch < 0: the chunk is the whole subset and its element count comes
  // from the rows; the voxel offsets are then computed from the sorted keys
  // (offsets_from_sorted) before the split.
  int rc;
  int* rc_cnt;
  int* pos;
  if ((rc = tmp.get(&rc_cnt, (size_t)n_rows_s + 1, kPoolRowCounts,
                    "row counts")) ||
      (rc = tmp.get(&pos, (size_t)n_rows_s + 1, kPoolRowPos, "row positions")))
    return rc;

  const dim3 grid_rows(blocks_for((long long)max_rows * kTrLanes), vpr);
  FIR_HIP_CHECK(hipMemsetAsync(rc_cnt, 0, sizeof(int) * ((size_t)n_rows_s + 1),
                               c->stream));
  const XY xy = cache_xy(cache);
  const int lb = (int)len_bytes(rm->fmt);
  hipLaunchKernelGGL(k_tr_count, grid_rows, dim3(kBlock), 0, c->stream, sr,
                     d_prefix, rm->d_off, rm->d_idx, rm->d_len, lb, rm->d_yrow,
                     xy, v0, v1, rc_cnt);
  FIR_CHECK_LAUNCH();
  size_t tb = 0;
  FIR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, rc_cnt, pos,
                                                 n_rows_s + 1, c->stream));
  unsigned char* scan_tmp;
  if ((rc = tmp.get(&scan_tmp, tb, kPoolRowScan, "scan scratch"))) return rc;
  FIR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb, rc_cnt, pos,
                                                 n_rows_s + 1, c->stream));
  int counted = -1;
  FIR_HIP_CHECK(hipMemcpyAsync(&counted, pos + n_rows_s, sizeof(int),
                               hipMemcpyDeviceToHost, c->stream));
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (n_ch < 0) n_ch = counted;
  if (counted < 0 || (long long)counted != n_ch)
    return fail(FIR_E_STATE, "path cache: voxels [%d, %d) hold %lld elements "
                             "by histogram, %d by rows", v0, v1, n_ch, counted);
  unsigned *k0, *k1;
  unsigned long long *val0, *val1;
  if ((rc = tmp.get(&k0, n_ch, kPoolKeys0, "sort keys")) ||
      (rc = tmp.get(&k1, n_ch, kPoolKeys1, "sort keys")) ||
      (rc = tmp.get(&val0, n_ch, kPoolVals0, "sort values")) ||
      (rc = tmp.get(&val1, n_ch, kPoolVals1, "sort values")))
    return rc;
  if (n_ch == 0)
    return offsets_from_sorted ? offsets_from_sorted(k0, 0) : FIR_OK;
  hipLaunchKernelGGL(k_tr_expand<F>, grid_rows, dim3(kBlock), 0, c->stream, sr,
                     d_prefix, rm->d_off, rm->d_idx, rm->d_yrow, xy,
                     (const typename Len<F>::T*)rm->d_len, v0, v1, pos, n_ch,
                     k0, val0);
  FIR_CHECK_LAUNCH();

  // Stable radix sort by voxel: elements of one voxel keep their row order.
  int bits = 1;
  while ((1LL << bits) < (long long)(v1 - v0)) ++bits;
  hipcub::DoubleBuffer<unsigned> keys(k0, k1);
  hipcub::DoubleBuffer<unsigned long long> vals(val0, val1);
  size_t sb = 0;
  FIR_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
    nullptr, sb, keys, vals, (int)n_ch, 0, bits, c->stream));
  unsigned char* sort_tmp;
  if ((rc = tmp.get(&sort_tmp, sb, kPoolSort, "sort scratch"))) return rc;
  FIR_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
    sort_tmp, sb, keys, vals, (int)n_ch, 0, bits, c->stream));
  if (offsets_from_sorted &&
      (rc = offsets_from_sorted(keys.Current(), n_ch)))
    return rc;
  hipLaunchKernelGGL(k_tr_split<F>, dim3(ew_blocks(n_ch)), dim3(kBlock), 0,
                     c->stream, keys.Current(), vals.Current(), n_ch, v0,
                     voff_p, voff_u, cache->bp_lanes, cache->d_trow,
                     (typename Len<F>::T*)cache->d_tlen);
  FIR_CHECK_LAUNCH();
  FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
  return FIR_OK;
}

int build_transpose(fir_ctx* c, fir_cache* cache)
{
  RowMatrix* rm = cache->rm.get();
  const int S = cache->n_subsets;
  const int K = cache->n_cached;
  const long long nvox = cache->nvox;
  int rc;
  // Padding adds at most 3 slots per (subset, voxel).
  const long long slots = rm->n_elems + 3LL * K * nvox;
  if ((rc = dev_alloc(&cache->d_voff, (size_t)K * (nvox + 1), "voxel offsets")))
    return rc;
  if ((rc = dev_alloc(&cache->d_trow, slots, "transposed rows"))) return rc;
  if ((rc = dev_alloc((unsigned char**)&cache->d_tlen,
                      slots * len_bytes(rm->fmt), "transposed lengths")))
    return rc;
  FIR_HIP_CHECK(hipMemsetAsync(cache->d_trow, 0, sizeof(int) * slots,
                               c->stream));
  FIR_HIP_CHECK(hipMemsetAsync(cache->d_tlen, 0, slots * len_bytes(rm->fmt),
                               c->stream));
  cache->t_slots = slots;

  std::vector<long long> voff(nvox + 1);
  long long base = 0, real = 0;
  Pool tmp;
  // The largest sort any subset needs, reserved up front so that the
  // buffers are allocated once.
  {
    long long max_slots = 0;
    int max_rows_s = 0;
    for (int s = 0; s < K; ++s)
    {
      long long sl = 0;
      int nr = 0;
      for (int k = 0; k < fir::views_in_subset(rm->n_views, S, s); ++k)
      {
        const int view = k * S + s;
        sl += rm->view_elem[view + 1] - rm->view_elem[view];
        nr += rm->view_row[view + 1] - rm->view_row[view];
      }
      max_slots = std::max(max_slots, sl);
      max_rows_s = std::max(max_rows_s, nr);
    }
    const size_t n_sort = (size_t)std::min(max_slots, c->cache_chunk);
    unsigned* ku;
    unsigned long long* vu;
    int* ri;
    if ((rc = tmp.get(&ri, (size_t)max_rows_s + 1, kPoolRowCounts,
                      "row counts")) ||
        (rc = tmp.get(&ri, (size_t)max_rows_s + 1, kPoolRowPos,
                      "row positions")) ||
        (rc = tmp.get(&ku, n_sort, kPoolKeys0, "sort keys")) ||
        (rc = tmp.get(&ku, n_sort, kPoolKeys1, "sort keys")) ||
        (rc = tmp.get(&vu, n_sort, kPoolVals0, "sort values")) ||
        (rc = tmp.get(&vu, n_sort, kPoolVals1, "sort values")))
      return rc;
  }
  for (int s = 0; s < K; ++s)
  {
    const int vpr = fir::views_in_subset(rm->n_views, S, s);
    std::vector<int> prefix(vpr + 1, 0);
    int max_rows = 0;
    for (int k = 0; k < vpr; ++k)
    {
      const int view = k * S + s;
      const int nr = rm->view_row[view + 1] - rm->view_row[view];
      prefix[k + 1] = prefix[k] + nr;
      max_rows = std::max(max_rows, nr);
    }
    const int n_rows_s = prefix[vpr];
    int* d_prefix;
    if ((rc = tmp.get(&d_prefix, vpr + 1, kPoolPrefix, "subset row prefix")))
      return rc;
    FIR_HIP_CHECK(hipMemcpyAsync(d_prefix, prefix.data(), sizeof(int) * (vpr + 1),
                                 hipMemcpyHostToDevice, c->stream));

    // Voxel counts of the subset's (non-zero-length) elements give the
    // padded offsets (stored) and unpadded offsets (chunk planning, ranks).
    // A subset of at most cache_chunk slots is one chunk: its counts come
    // from the sorted keys; larger subsets take a histogram first.
    unsigned* cnt;
    long long* voff_u;
    if ((rc = tmp.get(&cnt, nvox + 1, kPoolVoxCounts, "voxel counts")) ||
        (rc = tmp.get(&voff_u, nvox + 1, kPoolVoxOffsets, "voxel offsets")))
      return rc;
    long long* voff_s = cache->d_voff + (long long)s * (nvox + 1);
    long long end_p = 0;
    auto offsets = [&]() -> int {
      hipcub::TransformInputIterator<long long, Pad4, const unsigned*> cnt_p(
        cnt, Pad4());
      hipcub::TransformInputIterator<long long, ToI64, const unsigned*> cnt_u(
        cnt, ToI64());
      size_t tb = 0, tb2 = 0;
      FIR_HIP_CHECK(hipcub::DeviceScan::ExclusiveScan(
        nullptr, tb, cnt_p, voff_s, hipcub::Sum(), base, (int)(nvox + 1),
        c->stream));
      FIR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(
        nullptr, tb2, cnt_u, voff_u, (int)(nvox + 1), c->stream));
      unsigned char* scan_tmp;
      int r2;
      if ((r2 = tmp.get(&scan_tmp, std::max(tb, tb2), kPoolVoxScan,
                        "scan scratch")))
        return r2;
      FIR_HIP_CHECK(hipcub::DeviceScan::ExclusiveScan(
        scan_tmp, tb, cnt_p, voff_s, hipcub::Sum(), base, (int)(nvox + 1),
        c->stream));
      FIR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(
        scan_tmp, tb2, cnt_u, voff_u, (int)(nvox + 1), c->stream));
      FIR_HIP_CHECK(hipMemcpyAsync(&end_p, voff_s + nvox, sizeof(long long),
                                   hipMemcpyDeviceToHost, c->stream));
      FIR_HIP_CHECK(hipMemcpyAsync(voff.data(), voff_u,
                                   sizeof(long long) * (nvox + 1),
                                   hipMemcpyDeviceToHost, c->stream));
      FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
      if (end_p > slots)
        return fail(FIR_E_STATE, "path cache: transposed slots overflow");
      return FIR_OK;
    };
    long long sub_slots = 0;
    for (int k = 0; k < vpr; ++k)
    {
      const int view = k * S + s;
      sub_slots += rm->view_elem[view + 1] - rm->view_elem[view];
    }
    const SubsetRows sr{rm->d_view_row, S, s, 0};
    auto chunk = [&](long long v0, long long v1, long long n_ch,
                     const std::function<int(const unsigned*, long long)>&
                       from_sorted) -> int {
      switch (rm->fmt)
      {
      case FIR_LEN_SNORM16:
        return transpose_chunk<FIR_LEN_SNORM16>(
          c, rm, cache, sr, d_prefix, vpr, max_rows, n_rows_s, (int)v0,
          (int)v1, voff_s, voff_u, n_ch, from_sorted, tmp);
      case FIR_LEN_FP16:
        return transpose_chunk<FIR_LEN_FP16>(
          c, rm, cache, sr, d_prefix, vpr, max_rows, n_rows_s, (int)v0,
          (int)v1, voff_s, voff_u, n_ch, from_sorted, tmp);
      default:
        return transpose_chunk<FIR_LEN_FP32>(
          c, rm, cache, sr, d_prefix, vpr, max_rows, n_rows_s, (int)v0,
          (int)v1, voff_s, voff_u, n_ch, from_sorted, tmp);
      }
    };
    const long long chunk_max = c->cache_chunk;
    if (n_rows_s > 0 && sub_slots <= chunk_max)
    {
      rc = chunk(0, nvox, -1, [&](const unsigned* keys, long long n) -> int {
        hipLaunchKernelGGL(k_tr_counts_sorted, dim3(blocks_for(nvox + 1)),
                           dim3(kBlock), 0, c->stream, keys, n, (int)nvox,
                           cnt);
        FIR_CHECK_LAUNCH();
        return offsets();
      });
      if (rc) return rc;
    }
    else
    {
      FIR_HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned) * (nvox + 1),
                                   c->stream));
      if (max_rows > 0)
      {
        hipLaunchKernelGGL(k_tr_hist,
                           dim3(blocks_for((long long)max_rows * kTrLanes), vpr),
                           dim3(kBlock), 0, c->stream, sr, rm->d_off,
                           rm->d_idx, rm->d_len, (int)len_bytes(rm->fmt),
                           rm->d_yrow, cache_xy(cache), cnt);
        FIR_CHECK_LAUNCH();
      }
      if ((rc = offsets())) return rc;
      long long v0 = 0;
      while (v0 < nvox && n_rows_s > 0)
      {
        // Largest voxel range starting at v0 with at most chunk_max
        // elements.
        const long long limit = voff[v0] + chunk_max;
        long long v1 =
          std::upper_bound(voff.begin() + v0 + 1, voff.end(), limit) -
          voff.begin() - 1;
        if (v1 <= v0)
          return fail(FIR_E_STATE, "path cache: voxel %lld holds more than "
                                   "%lld elements", v0, chunk_max);
        if (v1 > nvox) v1 = nvox;
        const long long n_ch = voff[v1] - voff[v0];
        if (n_ch > 0 && (rc = chunk(v0, v1, n_ch, {}))) return rc;
        v0 = v1;
      }
    }
    base = end_p;
    real += voff[nvox];
  }
  if (real > rm->n_elems)
    return fail(FIR_E_STATE, "path cache: transposed %lld of %lld elements",
                real, rm->n_elems);
  return FIR_OK;
}

int check_cache(const fir_cache* cache)
{
  if (!cache || !cache->ctx || !cache->rm)
    return fail(FIR_E_ARG, "path cache is NULL");
  return FIR_OK;
}

RowData row_data(const RowMatrix* rm)
{
  return RowData{rm->d_off, rm->d_idx, rm->d_len, rm->d_bin, rm->d_factor,
                 rm->d_yrow};
}

int max_rows_of_subset(const RowMatrix* rm, int S, int s)
{
  int m = 0;
  const int vpr = fir::views_in_subset(rm->n_views, S, s);
  for (int k = 0; k < vpr; ++k)
  {
    const int view = k * S + s;
    m = std::max(m, rm->view_row[view + 1] - rm->view_row[view]);
  }
  return m;
}

// Blocks of a projector launch: a multiple of 8 with the XCD mapping.
unsigned grid_blocks(const fir_cache* cache, long long threads)
{
  unsigned b = blocks_for(threads);
  if (cache->xcd_map) b = (b + 7u) & ~7u;
  return b;
}

// The y-fastest copy (and, for nb > 1 or interleaved slices, the
// x-fastest copy) of the images of one launch.
void pack_images(fir_cache* cache, int wd, const ImgIn& in, int nb)
{
  const XY xy = cache_xy(cache);
  int zq_shift = 0;
  while ((1 << zq_shift) < xy.zq) ++zq_shift;
  const int na = std::max(xy.nx, xy.ny);
  const dim3 grid(blocks_for((long long)na * xy.zq * wd),
                  (unsigned)(xy.nx + xy.ny),
                  (unsigned)((cache->vol.nz + xy.zq - 1) / xy.zq));
  hipStream_t st = cache->ctx->stream;
  float* X = wd > 1 || xy.zq > 1 ? cache->d_x : nullptr;
  const int nz = cache->vol.nz;
  switch (wd)
  {
  case 1: hipLaunchKernelGGL(k_pc_pack<1>, grid, dim3(kBlock), 0, st, X, cache->d_xy, in, nb, xy, zq_shift, nz); break;
  case 2: hipLaunchKernelGGL(k_pc_pack<2>, grid, dim3(kBlock), 0, st, X, cache->d_xy, in, nb, xy, zq_shift, nz); break;
  case 4: hipLaunchKernelGGL(k_pc_pack<4>, grid, dim3(kBlock), 0, st, X, cache->d_xy, in, nb, xy, zq_shift, nz); break;
  default: hipLaunchKernelGGL(k_pc_pack<8>, grid, dim3(kBlock), 0, st, X, cache->d_xy, in, nb, xy, zq_shift, nz); break;
  }
}

template <int F>
void launch_forward(fir_cache* cache, int S, int s, const float* img,
                    float* sino)
{
  const RowMatrix* rm = cache->rm.get();
  const int lanes = cache->fwd_lanes;
  const int vpr = fir::views_in_subset(rm->n_views, S, s);
  const long long threads = (long long)max_rows_of_subset(rm, S, s) * lanes;
  const dim3 grid(grid_blocks(cache, std::max(threads, 1LL)), vpr);
  hipLaunchKernelGGL((k_pc_forward<F>), grid, dim3(kBlock), 0,
                     cache->ctx->stream,
                     SubsetRows{rm->d_view_row, S, s, cache->xcd_map},
                     row_data(rm), lanes, img, cache->d_xy, sino);
}

template <int F, int NB>
void launch_bp_nb(fir_cache* cache, int s, const float* w, const VoxOut& out)
{
  const int lanes = cache->bp_lanes;
  const long long threads = cache->nvox * lanes;
  const dim3 grid(grid_blocks(cache, threads));
  const long long* voff = cache->d_voff + (long long)s * (cache->nvox + 1);
  hipLaunchKernelGGL((k_pc_backproject<F, NB>), grid, dim3(kBlock), 0,
                     cache->ctx->stream, voff, cache->d_trow, cache->d_tlen,
                     cache->nvox, lanes, cache->xcd_map,
                     cache->zfast && NB <= 2 ? cache->vol.nz : 0, w, out);
}

template <int F>
void launch_backproject(fir_cache* cache, int s, int nb, const float* w,
                        const VoxOut& out)
{
  switch (nb)
  {
  case 1: launch_bp_nb<F, 1>(cache, s, w, out); break;
  case 2: launch_bp_nb<F, 2>(cache, s, w, out); break;
  case 4: launch_bp_nb<F, 4>(cache, s, w, out); break;
  default: launch_bp_nb<F, 8>(cache, s, w, out); break;
  }
}

template <int F, int NB>
void launch_rows_nb(fir_cache* cache, int s, const float* X,
                    const OsemRowsIO& io, float* w)
{
  const RowMatrix* rm = cache->rm.get();
  const int S = cache->n_subsets;
  const int lanes = cache->fwd_lanes;
  const int vpr = fir::views_in_subset(rm->n_views, S, s);
  const long long threads = (long long)max_rows_of_subset(rm, S, s) * lanes;
  const dim3 grid(grid_blocks(cache, std::max(threads, 1LL)), vpr);
  hipLaunchKernelGGL((k_pc_osem_rows<F, NB>), grid, dim3(kBlock), 0,
                     cache->ctx->stream,
                     SubsetRows{rm->d_view_row, S, s, cache->xcd_map},
                     row_data(rm), lanes, X, cache->d_xy, io, w);
}

// Whether one-image OSEM passes of this cache take the sheet form.
bool use_sheets(const fir_cache* cache)
{
  return cache->sheets && cache->rm->d_sel;
}

// Subset s's families for the sheet kernels: their order table and count.
const int* family_order(const fir_cache* cache, int s, int* n)
{
  *n = cache->ford_first[s + 1] - cache->ford_first[s];
  return cache->d_fam_ord + cache->ford_first[s];
}

// k_pc_osem_rows over the sheet form for one image (img x fastest).
template <int F>
void launch_rows_sheet(fir_cache* cache, int s, const float* img,
                       const OsemRowsIO& io, float* w)
{
  const RowMatrix* rm = cache->rm.get();
  int nf;
  const int* ford = family_order(cache, s, &nf);
  if (nf == 0) return;
  hipLaunchKernelGGL((k_pc_osem_rows_sheet<F>), dim3(nf), dim3(kSheetBlock),
                     0, cache->ctx->stream, ford, rm->d_fam_row,
                     rm->d_fam_off, rm->d_fam_xy, rm->d_fam_zr, rm->d_sel,
                     row_data(rm), cache->fwd_lanes, img,
                     cache->vol.nx * cache->vol.ny, cache->vol.nz, io, w);
}

// Whether one-image OSEM passes of this cache backproject family by family.
bool use_family_bp(const fir_cache* cache)
{
  return use_sheets(cache) && cache->family_bp && cache->d_part;
}

// One-image OSEM pass family by family: line integrals and weights from the
// LDS-staged sheets, each family's backprojection into its partial sheet,
// then the per-voxel sums of the partials (overwriting bp).
template <int F>
void launch_osem_family(fir_cache* cache, int s, const float* img,
                        const OsemRowsIO& io, float* w, float* bp)
{
  const RowMatrix* rm = cache->rm.get();
  const int nxy = cache->vol.nx * cache->vol.ny, nz = cache->vol.nz;
  int nf;
  const int* ford = family_order(cache, s, &nf);
  if (nf > 0)
    hipLaunchKernelGGL((k_pc_osem_fam<F>), dim3(nf), dim3(kSheetBlock), 0,
                       cache->ctx->stream, ford, rm->d_fam_row,
                       rm->d_fam_off, rm->d_fam_xy, rm->d_fam_zr, rm->d_sel,
                       row_data(rm), cache->fwd_lanes, img, nxy, nz, io, w,
                       rm->d_fl_slot, rm->d_fl_gb, rm->d_fl, rm->d_fl_nact,
                       cache->d_fam_loc, cache->d_part);
  VoxOut out;
  std::memset(&out, 0, sizeof(out));
  out.p[0] = bp;
  const size_t tile = red_tile_bytes(1, nz);
  hipLaunchKernelGGL(k_pc_fam_reduce<1>,
                     dim3((nxy + kRedBlock / 64 - 1) / (kRedBlock / 64)),
                     dim3(kRedBlock), tile,
                     cache->ctx->stream,
                     cache->d_fr_off + (long long)s * (nxy + 1),
                     cache->d_fr_ent, cache->d_part, nxy, nz,
                     out, tile > 0 ? 1 : 0);
}

// Family backprojection of nb images (nb = 1, 2, 4 or 8) of subset s with
// per-row weights w[r * nb + b], into out.p[b] (overwritten).
template <int F, int NB>
void launch_fam_bp_nb(fir_cache* cache, int s, const float* w,
                      const VoxOut& out)
{
  const RowMatrix* rm = cache->rm.get();
  const int nxy = cache->vol.nx * cache->vol.ny, nz = cache->vol.nz;
  int nf;
  const int* ford = family_order(cache, s, &nf);
  if (nf > 0)
    hipLaunchKernelGGL((k_pc_fam_bp<F, NB>), dim3(nf), dim3(kSheetBlock), 0,
                       cache->ctx->stream, ford, rm->d_fam_row,
                       rm->d_fam_off, cache->d_fam_loc, nz, w, rm->d_fl_slot,
                       rm->d_fl_gb, rm->d_fl, rm->d_fl_nact, cache->d_part,
                       rm->d_fam_zr);
  const size_t tile = red_tile_bytes(NB, nz);
  hipLaunchKernelGGL(k_pc_fam_reduce<NB>,
                     dim3((nxy + kRedBlock / 64 - 1) / (kRedBlock / 64)),
                     dim3(kRedBlock), tile,
                     cache->ctx->stream,
                     cache->d_fr_off + (long long)s * (nxy + 1),
                     cache->d_fr_ent, cache->d_part, nxy, nz,
                     out, tile > 0 ? 1 : 0);
}

template <int F>
void launch_fam_bp(fir_cache* cache, int s, int nb, const float* w,
                   const VoxOut& out)
{
  switch (nb)
  {
  case 1: launch_fam_bp_nb<F, 1>(cache, s, w, out); break;
  case 2: launch_fam_bp_nb<F, 2>(cache, s, w, out); break;
  case 4: launch_fam_bp_nb<F, 4>(cache, s, w, out); break;
  default: launch_fam_bp_nb<F, 8>(cache, s, w, out); break;
  }
}

// Rows per thread of the slab kernel's instances.
constexpr int kSlabRounds[] = {1, 2, 3, 5};

// Slab launch shape for a cache: rows per thread R (an instance) and the
// block size (lanes * rows per round, whole waves), or R = 0 when a
// family's rows do not fit in kSlabBlockMax threads times 5 rounds.
struct SlabShape
{
  int R = 0;
  int block = 0;
  int slab_px = 0;
  int buf_f4 = 0;  // float4 per LDS buffer (whole waves' worth)
  size_t lds = 0;  // bytes of the two buffers
};

// Bytes of one of the two LDS slab buffers: FIR_SLAB_KB (8 to 80),
// default 40 (two blocks of 4-lane rows fill a CU's 160 KB; 16 to 40 KB
// measured within 10% of each other at 160 x 160 x 47 with 8 images).
int slab_bytes()
{
  static const int kb = [] {
    const char* e = std::getenv("FIR_SLAB_KB");
    const int v = e ? std::atoi(e) : 0;
    return v >= 8 && v <= 80 ? v : 40;
  }();
  return kb * 1024;
}

SlabShape slab_shape(const fir_cache* cache, int nb)
{
  SlabShape sh;
  const RowMatrix* rm = cache->rm.get();
  const int lanes = cache->fwd_lanes;
  const int rows = rm->n_pairs;
  for (int R : kSlabRounds)
  {
    const int per_round = (rows + R - 1) / R;
    const int block = (per_round * lanes + 63) / 64 * 64;
    if (block <= kSlabBlockMax)
    {
      sh.R = R;
      sh.block = block;
      break;
    }
  }
  const int n4 = cache->vol.nz * nb / 4;
  sh.slab_px = slab_bytes() / (n4 * 16);
  sh.buf_f4 = (sh.slab_px * n4 + 63) / 64 * 64;
  sh.lds = 2 * 16 * (size_t)sh.buf_f4;
  if (sh.slab_px < 1) sh.R = 0;
  return sh;
}

// Whether a batch of nb images (a launch width) takes the slab kernel.
bool use_slabs(const fir_cache* cache, int nb)
{
  const RowMatrix* rm = cache->rm.get();
  return (nb == 4 || nb == 8) && cache->slabs && use_family_bp(cache) &&
    rm->line_order && cache->fwd_lanes == rm->lanes &&
    slab_shape(cache, nb).R > 0;
}

// The pixel-major copies of a launch's images, in cache->d_x.
void pack_images_pz(fir_cache* cache, int wd, const ImgIn& in, int nb)
{
  const int nxy = cache->vol.nx * cache->vol.ny, nz = cache->vol.nz;
  const dim3 grid((nxy + kPzPx - 1) / kPzPx, (nz + kPzZ - 1) / kPzZ);
  hipStream_t st = cache->ctx->stream;
  if (wd == 8)
    hipLaunchKernelGGL(k_pc_pack_pz<8>, grid, dim3(256), 0, st, cache->d_x,
                       in, nb, nxy, nz);
  else if (wd == 4)
    hipLaunchKernelGGL(k_pc_pack_pz<4>, grid, dim3(256), 0, st, cache->d_x,
                       in, nb, nxy, nz);
  else
    hipLaunchKernelGGL(k_pc_pack_pz<1>, grid, dim3(256), 0, st, cache->d_x,
                       in, 1, nxy, nz);
}

template <int F, int NB, int R, int D>
void launch_rows_slab_r(fir_cache* cache, int s, const SlabShape& sh,
                        const OsemRowsIO& io, float* w)
{
  const RowMatrix* rm = cache->rm.get();
  const int nz = cache->vol.nz;
  int nf;
  const int* ford = family_order(cache, s, &nf);
  if (nf == 0) return;
  auto kern = k_pc_osem_rows_slab<F, NB, R, D>;
  if (sh.lds > 65536)
  {
    // (beyond the default dynamic LDS limit; set once per instance)
    static size_t allowed = 0;
    if (allowed < sh.lds &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sh.lds) == hipSuccess)
      allowed = sh.lds;
  }
  hipLaunchKernelGGL(kern, dim3(nf), dim3(sh.block), sh.lds,
                     cache->ctx->stream, ford,
                     rm->d_fam_row, rm->d_fam_off, rm->d_fam_xy, rm->d_sel,
                     row_data(rm), cache->fwd_lanes, cache->d_x, nz,
                     sh.slab_px, sh.buf_f4, io, w);
}

// (two quads of each row in hand: within 1-4% of one at 4 lanes)
template <int F, int NB>
void launch_rows_slab_nb(fir_cache* cache, int s, const SlabShape& sh,
                         const OsemRowsIO& io, float* w)
{
  switch (sh.R)
  {
  case 1: launch_rows_slab_r<F, NB, 1, 2>(cache, s, sh, io, w); break;
  case 2: launch_rows_slab_r<F, NB, 2, 2>(cache, s, sh, io, w); break;
  case 3: launch_rows_slab_r<F, NB, 3, 2>(cache, s, sh, io, w); break;
  default: launch_rows_slab_r<F, NB, 5, 2>(cache, s, sh, io, w); break;
  }
}

// Batch line integrals of wd = 4 or 8 images from slab-staged sheets (the
// images already packed pixel-major in cache->d_x).
template <int F>
void launch_rows_slab(fir_cache* cache, int s, int wd, const OsemRowsIO& io,
                      float* w)
{
  const SlabShape sh = slab_shape(cache, wd);
  if (wd == 8) launch_rows_slab_nb<F, 8>(cache, s, sh, io, w);
  else launch_rows_slab_nb<F, 4>(cache, s, sh, io, w);
}

template <int F>
void launch_osem_rows(fir_cache* cache, int s, int nb, const float* X,
                      const OsemRowsIO& io, float* w)
{
  switch (nb)
  {
  case 1: launch_rows_nb<F, 1>(cache, s, X, io, w); break;
  case 2: launch_rows_nb<F, 2>(cache, s, X, io, w); break;
  case 4: launch_rows_nb<F, 4>(cache, s, X, io, w); break;
  default: launch_rows_nb<F, 8>(cache, s, X, io, w); break;
  }
}

int batch_width(int n)
{
  return n <= 1 ? 1 : n <= 2 ? 2 : n <= 4 ? 4 : 8;
}

} // namespace

extern "C" {

int fir_set_cache_lanes(fir_ctx* c, int fwd_lanes, int bp_lanes)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  auto ok = [](int t) { return t >= 1 && t <= 64 && (t & (t - 1)) == 0; };
  if (!ok(fwd_lanes) || !ok(bp_lanes))
    return fail(FIR_E_ARG, "lanes per LOR / voxel must be a power of two "
                           "from 1 to 64");
  c->cache_lanes[0] = fwd_lanes;
  c->cache_lanes[1] = bp_lanes;
  return FIR_OK;
}

int fir_set_cache_sheets(fir_ctx* c, int enable)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  c->cache_sheets = enable ? 1 : 0;
  return FIR_OK;
}

int fir_set_cache_chunk(fir_ctx* c, int64_t max_elements)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (max_elements < 1024 || max_elements > (1LL << 30))
    return fail(FIR_E_ARG, "transpose chunk must hold 1024 to 2^30 elements");
  c->cache_chunk = max_elements;
  return FIR_OK;
}

int fir_set_cache_zslices(fir_ctx* c, int zq)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (zq < 1 || zq > 16 || (zq & (zq - 1)) != 0)
    return fail(FIR_E_ARG, "z-slices per layout group must be a power of two "
                           "from 1 to 16");
  c->cache_zq = zq;
  return FIR_OK;
}

} // extern "C"

// Shared body of fir_cache_create (plan NULL: every subset, the path matrix
// shared with the context's other caches) and fir_cache_create_partial.
static int cache_create(fir_ctx* c, const fir_vol_desc* v, int n_subsets,
                        int len_format, PartialPlan* plan, fir_cache** out)
{
  if (!out) return fail(FIR_E_ARG, "output pointer is NULL");
  *out = nullptr;
  int rc = need_lors(c);
  if (rc) return rc;
  if (len_format != FIR_LEN_SNORM16 && len_format != FIR_LEN_FP16 &&
      len_format != FIR_LEN_FP32)
    return fail(FIR_E_ARG, "unknown length format %d", len_format);
  Grid g;
  if ((rc = make_grid(c, v, &g))) return rc;
  if ((rc = check_subsets(c, n_subsets))) return rc;
  FIR_HIP_CHECK(hipSetDevice(c->device));

  std::shared_ptr<RowMatrix> rm;
  for (auto& wp : c->row_matrices)
  {
    if (plan) break;
    auto sp = wp.lock();
    if (sp && sp->generation == c->generation && same_vol(sp->vol, *v) &&
        sp->walk_mode == c->walk_mode && sp->pair_order == c->lor_order &&
        sp->tang_major == c->tang_major && sp->lanes == c->cache_lanes[0] &&
        sp->zq == c->cache_zq && sp->fmt == len_format &&
        sp->sheets_req == c->cache_sheets)
    {
      rm = sp;
      break;
    }
  }
  if (plan)
  {
    plan->S = n_subsets;
    plan->nvox = (long long)v->nx * v->ny * v->nz;
    plan->image_bytes = 2LL * 4 * kMaxCacheBatch *
      layout_voxels(XY{v->nx, v->ny, c->cache_zq}, v->nz);
    if ((rc = build_rows(c, v, len_format, &rm, plan))) return rc;
  }
  else if (!rm)
  {
    if ((rc = build_rows(c, v, len_format, &rm))) return rc;
    c->row_matrices.erase(
      std::remove_if(c->row_matrices.begin(), c->row_matrices.end(),
                     [](const std::weak_ptr<RowMatrix>& w) { return w.expired(); }),
      c->row_matrices.end());
    c->row_matrices.push_back(rm);
  }
  if (rm->n_rows >= 0x7fffffff)
    return fail(FIR_E_ARG, "path cache: more than 2^31-1 crossing LORs");

  auto* cache = new fir_cache();
  cache->ctx = c;
  cache->rm = rm;
  cache->vol = *v;
  cache->nvox = (long long)v->nx * v->ny * v->nz;
  cache->n_subsets = n_subsets;
  cache->n_cached = plan ? plan->n_cached : n_subsets;
  cache->fwd_lanes = c->cache_lanes[0];
  cache->bp_lanes = c->cache_lanes[1];
  size_t free_b = 0, total_b = 0;
  FIR_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  // With family lists every backprojection runs family by family and the
  // voxel-major transpose is built only if tuning asks for the other
  // paths (fir_cache_tune); without them it is built now.
  const bool lists = rm->d_fl != nullptr;
  const long long need =
    (lists ? 0
           : (rm->n_elems + 3LL * cache->n_cached * cache->nvox) *
               (4 + (long long)len_bytes(len_format)) +
             (long long)cache->n_cached * (cache->nvox + 1) * 8) +
    ((long long)rm->n_rows + 2 * cache->nvox) * kMaxCacheBatch * 4;
  if ((long long)free_b < need + (1LL << 30))
  {
    delete cache;
    return fail(FIR_E_NOMEM,
                "path cache: the transpose needs %.2f GB of device memory, "
                "%.2f GB free",
                need / 1e9, free_b / 1e9);
  }
  // Per-row weights and interleaved images for up to 8 images, allocated
  // once so that later calls (and graphs capturing them) never reallocate.
  cache->w_floats = (size_t)rm->n_rows * kMaxCacheBatch;
  cache->x_floats = (size_t)layout_voxels(cache_xy(cache), v->nz) *
    kMaxCacheBatch;
  if ((rc = build_family_order(c, cache)) ||
      (rc = build_family_bp(c, cache)) ||
      (!use_family_bp(cache) && (rc = build_transpose(c, cache))) ||
      (rc = dev_alloc(&cache->d_w, cache->w_floats, "row weights")) ||
      (rc = dev_alloc(&cache->d_x, cache->x_floats, "interleaved images")) ||
      (rc = dev_alloc(&cache->d_xy, cache->x_floats, "y-fastest images")))
  {
    (void)hipStreamSynchronize(c->stream);
    delete cache;
    return rc;
  }
  // the image copies' padding slices are never gathered; keep them zero
  if (hipMemsetAsync(cache->d_x, 0, cache->x_floats * sizeof(float),
                     c->stream) != hipSuccess ||
      hipMemsetAsync(cache->d_xy, 0, cache->x_floats * sizeof(float),
                     c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
  {
    delete cache;
    return fail(FIR_E_HIP, "path cache: clearing the image copies failed");
  }
  *out = cache;
  return FIR_OK;
}

extern "C" {

int fir_cache_create(fir_ctx* c, const fir_vol_desc* v, int n_subsets,
                     int len_format, fir_cache** out)
{
  return cache_create(c, v, n_subsets, len_format, nullptr, out);
}

int fir_cache_create_partial(fir_ctx* c, const fir_vol_desc* v, int n_subsets,
                             int len_format, int max_subsets,
                             int64_t max_bytes, fir_cache** out)
{
  if (out) *out = nullptr;
  if (max_subsets < 0) return fail(FIR_E_ARG, "max_subsets must be >= 0");
  PartialPlan plan;
  plan.max_cached = max_subsets;
  plan.max_bytes = max_bytes;
  return cache_create(c, v, n_subsets, len_format, &plan, out);
}

int fir_cache_subsets(const fir_cache* cache, int* n_subsets, int* n_cached)
{
  int rc = check_cache(cache);
  if (rc) return rc;
  if (n_subsets) *n_subsets = cache->n_subsets;
  if (n_cached) *n_cached = cache->n_cached;
  return FIR_OK;
}

int fir_cache_destroy(fir_cache* cache)
{
  if (!cache) return FIR_OK;
  if (cache->ctx)
  {
    (void)hipSetDevice(cache->ctx->device);
    (void)hipStreamSynchronize(cache->ctx->stream);
  }
  delete cache;
  return FIR_OK;
}

int fir_cache_info(const fir_cache* cache, int64_t* n_rows, int64_t* n_elements,
                   int64_t* device_bytes)
{
  int rc = check_cache(cache);
  if (rc) return rc;
  const RowMatrix* rm = cache->rm.get();
  if (n_rows) *n_rows = rm->n_rows;
  if (n_elements) *n_elements = rm->n_elems;
  if (device_bytes)
    *device_bytes = rm->bytes() +
      cache->t_slots * (4 + (long long)len_bytes(rm->fmt)) +
      (long long)cache->n_cached * (cache->nvox + 1) * 8 +
      (long long)(cache->w_floats + 2 * cache->x_floats) * 4 +
      cache->fr_bytes;
  return FIR_OK;
}

int fir_cache_forms(const fir_cache* cache, int* forms)
{
  int rc = check_cache(cache);
  if (rc) return rc;
  if (!forms) return fail(FIR_E_ARG, "NULL output");
  const RowMatrix* rm = cache->rm.get();
  *forms = (rm->d_sel ? FIR_FORM_SHEETS : 0) |
    (rm->d_sel && rm->d_fl && cache->d_part ? FIR_FORM_FAMILY_LISTS : 0) |
    (use_slabs(cache, kMaxCacheBatch) ? FIR_FORM_SLABS : 0);
  return FIR_OK;
}

int fir_cache_tune(fir_cache* cache, int fwd_lanes, int bp_lanes, int flags)
{
  int rc = check_cache(cache);
  if (rc) return rc;
  auto ok = [](int t) { return t >= 1 && t <= 64 && (t & (t - 1)) == 0; };
  if (!ok(fwd_lanes) || !ok(bp_lanes))
    return fail(FIR_E_ARG, "lanes per LOR / voxel must be a power of two "
                           "from 1 to 64");
  if (flags & ~(FIR_CACHE_XCD_MAP | FIR_CACHE_STORED_VOXELS |
                FIR_CACHE_NO_SHEETS | FIR_CACHE_NO_FAMILY_BP |
                FIR_CACHE_NO_SLABS))
    return fail(FIR_E_ARG, "unknown cache flags 0x%x", flags);
  const int sheets = cache->sheets, family_bp = cache->family_bp;
  cache->sheets = (flags & FIR_CACHE_NO_SHEETS) ? 0 : 1;
  cache->family_bp = (flags & FIR_CACHE_NO_FAMILY_BP) ? 0 : 1;
  if (!use_family_bp(cache) && !cache->d_voff)
  {
    // the transposed voxel lists, skipped at creation, are needed now; on
    // failure the cache keeps its previous settings
    FIR_HIP_CHECK(hipSetDevice(cache->ctx->device));
    if ((rc = build_transpose(cache->ctx, cache)))
    {
      (void)hipStreamSynchronize(cache->ctx->stream);
      (void)hipFree(cache->d_voff);
      (void)hipFree(cache->d_trow);
      (void)hipFree(cache->d_tlen);
      cache->d_voff = nullptr;
      cache->d_trow = nullptr;
      cache->d_tlen = nullptr;
      cache->t_slots = 0;
      cache->sheets = sheets;
      cache->family_bp = family_bp;
      return rc;
    }
  }
  cache->zfast = (flags & FIR_CACHE_STORED_VOXELS) ? 0 : 1;
  cache->slabs = (flags & FIR_CACHE_NO_SLABS) ? 0 : 1;
  cache->fwd_lanes = fwd_lanes;
  cache->bp_lanes = bp_lanes;
  cache->xcd_map = (flags & FIR_CACHE_XCD_MAP) ? 1 : 0;
  return FIR_OK;
}

int fir_cache_forward(fir_cache* cache, const float* img, float* sino)
{
  int rc = check_cache(cache);
  if (rc) return rc;
  if (!img || !sino) return fail(FIR_E_ARG, "NULL buffer");
  fir_ctx* c = cache->ctx;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  // LORs that miss the volume project to 0 (projections.cc:63-66); a
  // partial cache writes the bins of its subsets' views only.
  if (cache->n_cached == cache->n_subsets)
  {
    if ((rc = fir_fill(c, sino, c->n_bins, 0.0f))) return rc;
  }
  else
  {
    if (cache->rm->generation != c->generation)
      return fail(FIR_E_STATE, "path cache: the sinogram geometry changed "
                               "since the partial cache was built");
    hipLaunchKernelGGL(k_pc_zero_subset_bins, dim3(ew_blocks(c->n_bins)),
                       dim3(kBlock), 0, c->stream, sino, (long long)c->n_bins,
                       c->d_seg_base, c->d_n_axial, c->n_seg, c->n_tang,
                       cache->n_subsets, cache->n_cached);
    FIR_CHECK_LAUNCH();
  }
  ImgIn in;
  std::memset(&in, 0, sizeof(in));
  in.p[0] = img;
  pack_images(cache, 1, in, 1);
  FIR_CHECK_LAUNCH();
  const int fmt = cache->rm->fmt;
  // One launch over all views (S = 1, subset 0 covers every view).
  const float* X = cache->rm->zq > 1 ? cache->d_x : img;
  if (fmt == FIR_LEN_SNORM16)
    launch_forward<FIR_LEN_SNORM16>(cache, 1, 0, X, sino);
  else if (fmt == FIR_LEN_FP16)
    launch_forward<FIR_LEN_FP16>(cache, 1, 0, X, sino);
  else
    launch_forward<FIR_LEN_FP32>(cache, 1, 0, X, sino);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

static int cache_backward_impl(fir_cache* cache, const float* sino,
                               float* frames)
{
  int rc = check_cache(cache);
  if (rc) return rc;
  if (!frames) return fail(FIR_E_ARG, "NULL image");
  fir_ctx* c = cache->ctx;
  const RowMatrix* rm = cache->rm.get();
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_pc_row_weights, dim3(ew_blocks(rm->n_rows)),
                     dim3(kBlock), 0, c->stream, rm->d_bin, rm->d_factor,
                     rm->n_rows, sino, cache->d_w);
  FIR_CHECK_LAUNCH();
  for (int s = 0; s < cache->n_cached; ++s)
  {
    VoxOut out;
    std::memset(&out, 0, sizeof(out));
    out.p[0] = frames + (long long)s * cache->nvox;
    if (use_family_bp(cache))
    {
      if (rm->fmt == FIR_LEN_SNORM16)
        launch_fam_bp<FIR_LEN_SNORM16>(cache, s, 1, cache->d_w, out);
      else
        launch_fam_bp<FIR_LEN_FP16>(cache, s, 1, cache->d_w, out);
    }
    else if (rm->fmt == FIR_LEN_SNORM16)
      launch_backproject<FIR_LEN_SNORM16>(cache, s, 1, cache->d_w, out);
    else if (rm->fmt == FIR_LEN_FP16)
      launch_backproject<FIR_LEN_FP16>(cache, s, 1, cache->d_w, out);
    else
      launch_backproject<FIR_LEN_FP32>(cache, s, 1, cache->d_w, out);
    FIR_CHECK_LAUNCH();
  }
  return FIR_OK;
}

int fir_cache_backward(fir_cache* cache, const float* sino, float* frames)
{
  if (!sino) return fail(FIR_E_ARG, "NULL sinogram");
  return cache_backward_impl(cache, sino, frames);
}

int fir_cache_sensitivity(fir_cache* cache, float* frames)
{
  return cache_backward_impl(cache, nullptr, frames);
}

int fir_cache_osem_project(fir_cache* cache, int n_images,
                           const float* const* x, const float* const* y,
                           const float* const* bias, float* const* bp,
                           int subset)
{
  int rc = check_cache(cache);
  if (rc) return rc;
  if (n_images <= 0) return fail(FIR_E_ARG, "n_images must be > 0");
  if (!x || !y || !bp) return fail(FIR_E_ARG, "NULL pointer array");
  for (int i = 0; i < n_images; ++i)
    if (!x[i] || !y[i] || !bp[i]) return fail(FIR_E_ARG, "NULL buffer");
  if (subset < 0 || subset >= cache->n_subsets)
    return fail(FIR_E_ARG, "subset %d out of range", subset);
  if (subset >= cache->n_cached)
    return fail(FIR_E_ARG,
                "subset %d is not in the path cache (it holds subsets 0 to "
                "%d); project it with fir_osem_project",
                subset, cache->n_cached - 1);
  fir_ctx* c = cache->ctx;
  const RowMatrix* rm = cache->rm.get();
  FIR_HIP_CHECK(hipSetDevice(c->device));
  for (int i0 = 0; i0 < n_images; i0 += kMaxCacheBatch)
  {
    const int nb = std::min(kMaxCacheBatch, n_images - i0);
    const int wd = batch_width(nb);
    OsemRowsIO io;
    std::memset(&io, 0, sizeof(io));
    ImgIn in;
    std::memset(&in, 0, sizeof(in));
    VoxOut out;
    std::memset(&out, 0, sizeof(out));
    for (int b = 0; b < nb; ++b)
    {
      io.y[b] = y[i0 + b];
      io.bias[b] = bias ? bias[i0 + b] : nullptr;
      in.p[b] = x[i0 + b];
      out.p[b] = bp[i0 + b];
    }
    io.n_images = nb;
    // (the sheet form exists for the 16-bit length formats only)
    // one image: the sheets are staged from its pixel-major copy in d_x
    if (wd == 1 && use_family_bp(cache))
    {
      pack_images_pz(cache, 1, in, 1);
      FIR_CHECK_LAUNCH();
      // (the weights stay in LDS: no per-row copy is stored)
      if (rm->fmt == FIR_LEN_SNORM16)
        launch_osem_family<FIR_LEN_SNORM16>(cache, subset, cache->d_x, io,
                                            nullptr, bp[i0]);
      else
        launch_osem_family<FIR_LEN_FP16>(cache, subset, cache->d_x, io,
                                         nullptr, bp[i0]);
      FIR_CHECK_LAUNCH();
      continue;
    }
    if (wd == 1 && use_sheets(cache))
    {
      pack_images_pz(cache, 1, in, 1);
      FIR_CHECK_LAUNCH();
      if (rm->fmt == FIR_LEN_SNORM16)
      {
        launch_rows_sheet<FIR_LEN_SNORM16>(cache, subset, cache->d_x, io,
                                           cache->d_w);
        FIR_CHECK_LAUNCH();
        launch_backproject<FIR_LEN_SNORM16>(cache, subset, 1, cache->d_w, out);
      }
      else
      {
        launch_rows_sheet<FIR_LEN_FP16>(cache, subset, cache->d_x, io,
                                        cache->d_w);
        FIR_CHECK_LAUNCH();
        launch_backproject<FIR_LEN_FP16>(cache, subset, 1, cache->d_w, out);
      }
      FIR_CHECK_LAUNCH();
      continue;
    }
    if (use_slabs(cache, wd))
    {
      // batch line integrals from slab-staged sheets, then the family
      // backprojection of all its images from the weights
      pack_images_pz(cache, wd, in, nb);
      FIR_CHECK_LAUNCH();
      if (rm->fmt == FIR_LEN_SNORM16)
      {
        launch_rows_slab<FIR_LEN_SNORM16>(cache, subset, wd, io, cache->d_w);
        FIR_CHECK_LAUNCH();
        launch_fam_bp<FIR_LEN_SNORM16>(cache, subset, wd, cache->d_w, out);
      }
      else
      {
        launch_rows_slab<FIR_LEN_FP16>(cache, subset, wd, io, cache->d_w);
        FIR_CHECK_LAUNCH();
        launch_fam_bp<FIR_LEN_FP16>(cache, subset, wd, cache->d_w, out);
      }
      FIR_CHECK_LAUNCH();
      continue;
    }
    const float* X = wd > 1 || rm->zq > 1 ? cache->d_x : x[i0];
    pack_images(cache, wd, in, nb);
    FIR_CHECK_LAUNCH();
    if (wd > 1 && use_family_bp(cache))
    {
      // batch line integrals from the image copies, then the family
      // backprojection of all its images from the weights
      if (rm->fmt == FIR_LEN_SNORM16)
      {
        launch_osem_rows<FIR_LEN_SNORM16>(cache, subset, wd, X, io,
                                          cache->d_w);
        FIR_CHECK_LAUNCH();
        launch_fam_bp<FIR_LEN_SNORM16>(cache, subset, wd, cache->d_w, out);
      }
      else
      {
        launch_osem_rows<FIR_LEN_FP16>(cache, subset, wd, X, io, cache->d_w);
        FIR_CHECK_LAUNCH();
        launch_fam_bp<FIR_LEN_FP16>(cache, subset, wd, cache->d_w, out);
      }
      FIR_CHECK_LAUNCH();
      continue;
    }
    if (rm->fmt == FIR_LEN_SNORM16)
    {
      launch_osem_rows<FIR_LEN_SNORM16>(cache, subset, wd, X, io, cache->d_w);
      FIR_CHECK_LAUNCH();
      launch_backproject<FIR_LEN_SNORM16>(cache, subset, wd, cache->d_w, out);
    }
    else if (rm->fmt == FIR_LEN_FP16)
    {
      launch_osem_rows<FIR_LEN_FP16>(cache, subset, wd, X, io, cache->d_w);
      FIR_CHECK_LAUNCH();
      launch_backproject<FIR_LEN_FP16>(cache, subset, wd, cache->d_w, out);
    }
    else
    {
      launch_osem_rows<FIR_LEN_FP32>(cache, subset, wd, X, io, cache->d_w);
      FIR_CHECK_LAUNCH();
      launch_backproject<FIR_LEN_FP32>(cache, subset, wd, cache->d_w, out);
    }
    FIR_CHECK_LAUNCH();
  }
  return FIR_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Multi-job OSEM steps: many reconstructions, on several caches of one
// context, advanced by one sub-iteration each in one launch per phase.
// ---------------------------------------------------------------------------
struct fir_osem_step
{
  fir_ctx* ctx = nullptr;
  int fmt = 0;
  int n_jobs = 0;
  void* d_jobs = nullptr;      // StepJob[n_jobs]
  int2* d_items = nullptr;     // (job, family), the most work first: the
  int n_items = 0;             // jobs of 8-image slots, then of 4
  int n_items8 = 0;
  int* d_pack_pre = nullptr;   // [n_jobs + 1] first pack tile of each job
  int pack_blocks = 0;
  int* d_red_pre = nullptr;    // [n_jobs + 1] first reduction block
  int red_blocks = 0;
  firi::StepImage* d_imgs = nullptr;
  int n_imgs = 0;
  int img_tiles = 0;
  size_t red_tile = 0;         // LDS of the reduction's output tile (0: none)
  bool vec_gauss = false;      // the vectorised Gaussian passes
  int vtiles[3] = {0, 0, 0};   // their blocks over all images
  int vh[3] = {0, 0, 0};       // the half-widths every image shares
  int slab_R = 0, slab_block = 0;
  size_t slab_lds = 0;
  int grid_slab[2] = {0, 0}, grid_bp[2] = {0, 0};
  int* d_ctr = nullptr;        // work-queue counters of the persistent
                               // launches: next item, blocks done (each)
  bool gauss = false;
  float cut = 0.0f;
  int restore = 0;
  double restore_r = 0.0;
  ~fir_osem_step()
  {
    if (ctx) (void)hipSetDevice(ctx->device);
    (void)hipFree(d_jobs);
    (void)hipFree(d_items);
    (void)hipFree(d_pack_pre);
    (void)hipFree(d_red_pre);
    (void)hipFree(d_imgs);
    (void)hipFree(d_ctr);
  }
};

namespace
{

// One job of a step, as its kernels read it.
struct StepJob
{
  const int* fam_row;
  const int* fam_off;
  const int* fam_xy;
  const unsigned* sel;
  RowData rd;
  const unsigned* fl_slot;
  const unsigned* fl_gb;
  const unsigned* fl;
  const int* fl_nact;
  const int* fam_loc;
  const long long* fr_off; // the subset's per-pixel table offsets
  const int2* fr_ent;
  float* w;                // workspace: per-row weights of 8 images
  float* X;                // pixel-major copies of the images
  float* part;             // partial sheets
  OsemRowsIO io;
  ImgIn in;
  VoxOut out;
  int nxy, nz, lanes, slab_px, buf_f4, pack_tiles_x;
  int nb; // images per launch slot: 4 (jobs of up to 4 images) or 8
};

// The job of block b: the last job whose first block (pre) is <= b.
__device__ __forceinline__ int step_job_of(const int* __restrict__ pre,
                                           int n_jobs, int b)
{
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Pixel-major copies of every job's images (k_pc_pack_pz, 8 images).
__global__ __launch_bounds__(256) void k_step_pack(
  const StepJob* __restrict__ jobs, const int* __restrict__ pre, int n_jobs)
{
  __shared__ float tile[kPzZ * kMaxCacheBatch * kPzPitch];
  const int b = blockIdx.x;
  const int j = step_job_of(pre, n_jobs, b);
  const StepJob& J = jobs[j];
  const int t = b - pre[j];
  const int bz = t / J.pack_tiles_x, bx = t - bz * J.pack_tiles_x;
  if (J.nb == 4)
    pack_pz_tile<4>(J.X, J.in, J.io.n_images, J.nxy, J.nz, bx, bz, tile);
  else
    pack_pz_tile<8>(J.X, J.in, J.io.n_images, J.nxy, J.nz, bx, bz, tile);
}

// The next work item of a persistent kernel's block: a block-wide grab from
// the launch's queue counter q[0]; past the last item the block counts
// itself done in q[1], and the last block done resets both for the next
// launch (every block's last grab precedes its count).
__device__ __forceinline__ int step_next_item(int* q, int* s_it, int n_items)
{
  if (threadIdx.x == 0) *s_it = atomicAdd(q, 1);
  __syncthreads();
  // (wave-uniform: the job's pointers are then scalar loads)
  const int it = __builtin_amdgcn_readfirstlane(*s_it);
  if (it >= n_items && threadIdx.x == 0 &&
      atomicAdd(q + 1, 1) == (int)gridDim.x - 1)
  {
    atomicExch(q, 0);
    atomicExch(q + 1, 0);
  }
  return it;
}

// Persistent: the blocks of a grid that fills the device take the
// (job, family) items of the jobs of NB-image slots from a queue, in the
// table's order (the most work first, so the launch ends on small
// families). Each item is one family of k_pc_osem_rows_slab. Registers
// are held to two 768-thread blocks per CU, as that kernel's.
template <int F, int NB, int R, int D>
__global__ __launch_bounds__(kSlabBlockMax)
__attribute__((amdgpu_waves_per_eu(6))) void k_step_slab(
  const StepJob* __restrict__ jobs, const int2* __restrict__ items,
  int n_items, int* __restrict__ q)
{
  extern __shared__ float4 slab[];
  // the grabbed item is passed through the (then idle) slab buffer: a
  // static LDS word beside two 40 KB buffers would leave room for one
  // block per CU instead of two
  int* s_it = reinterpret_cast<int*>(slab);
  for (;;)
  {
    const int it = step_next_item(q, s_it, n_items);
    __syncthreads(); // every wave has the item before the staging reuses it
    if (it >= n_items) break;
    const int2 jf = items[it];
    const StepJob& J = jobs[jf.x];
    slab_family<F, NB, R, D>(jf.y, slab, J.fam_row, J.fam_off, J.fam_xy,
                             J.sel, J.rd, J.lanes, J.X, J.nz, J.slab_px,
                             J.buf_f4, J.io, J.w);
    __syncthreads(); // the next item's staging reuses the buffers
  }
}

// Persistent as k_step_slab: each item one family of k_pc_fam_bp (the
// weights the slab phase wrote); two 1024-thread blocks per CU.
template <int F, int NB>
__global__ __launch_bounds__(kSheetBlock)
__attribute__((amdgpu_waves_per_eu(8))) void k_step_fam_bp(
  const StepJob* __restrict__ jobs, const int2* __restrict__ items,
  int n_items, int* __restrict__ q)
{
  __shared__ float wr[kFamRowsMax * NB];
  __shared__ int s_it;
  for (;;)
  {
    const int it = step_next_item(q, &s_it, n_items);
    if (it >= n_items) break;
    const int2 jf = items[it];
    const StepJob& J = jobs[jf.x];
    fam_bp_family<F, NB>(jf.y, wr, J.fam_row, J.fam_off, J.fam_loc, J.nz,
                         J.w, J.fl_slot, J.fl_gb, J.fl, J.fl_nact, J.part);
    __syncthreads(); // the next item overwrites the weights
  }
}

// Per-voxel sums of every job's partial sheets into its bp images
// (k_pc_fam_reduce, 8 images), one wave per pixel.
__global__ __launch_bounds__(kRedBlock) void k_step_reduce(
  const StepJob* __restrict__ jobs, const int* __restrict__ pre, int n_jobs,
  int stage)
{
  extern __shared__ float red_tile[];
  const int b = blockIdx.x;
  const int j = step_job_of(pre, n_jobs, b);
  const StepJob& J = jobs[j];
  const int w = (int)threadIdx.x >> 6;
  const int px0 = (b - pre[j]) * kRedPx;
  const int px = __builtin_amdgcn_readfirstlane(px0 + w);
  float* tile = stage ? red_tile : nullptr;
  if (px < J.nxy)
  {
    if (J.nb == 4)
      fam_reduce_pixel<4>(px, J.fr_off, J.fr_ent, J.part, J.nxy, J.nz, J.out,
                          tile, w);
    else
      fam_reduce_pixel<8>(px, J.fr_off, J.fr_ent, J.part, J.nxy, J.nz, J.out,
                          tile, w);
  }
  if (stage)
  {
    __syncthreads();
    if (J.nb == 4)
      fam_reduce_store<4>(red_tile, px0, J.nxy, J.nz, J.out);
    else
      fam_reduce_store<8>(red_tile, px0, J.nxy, J.nz, J.out);
  }
}

constexpr size_t kWsAlign = 256;

size_t ws_round(size_t bytes) { return (bytes + kWsAlign - 1) / kWsAlign * kWsAlign; }

// Workspace layout of a job on a cache: weights, pixel-major images,
// partial sheets, the Gaussian's two intermediate images per image.
struct WsLayout
{
  size_t w = 0, x = 0, part = 0, t1 = 0, t2 = 0, total = 0;
};

WsLayout ws_layout(const fir_cache* cache)
{
  WsLayout L;
  const size_t nvox = (size_t)cache->nvox;
  size_t o = 0;
  L.w = o;
  o += ws_round(4 * (size_t)cache->rm->n_rows * kMaxCacheBatch);
  L.x = o;
  o += ws_round(4 * nvox * kMaxCacheBatch);
  L.part = o;
  o += ws_round(4 * (size_t)std::max(cache->part_floats, 1LL));
  L.t1 = o;
  o += ws_round(4 * nvox * kMaxCacheBatch);
  L.t2 = o;
  o += ws_round(4 * nvox * kMaxCacheBatch);
  L.total = o;
  return L;
}

// Blocks of a kernel that fit on the device at once (the persistent grid):
// per CU as many as its 160 KB of LDS and 32 wave slots hold (the kernels
// keep their registers within that); the runtime's occupancy figure when
// it allows more. A larger grid is still correct, its extra blocks taking
// their items once others finish.
int resident_blocks(const void* kernel, int block, size_t lds, int device)
{
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block,
                                                   lds) != hipSuccess)
    per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            device) != hipSuccess)
    return 0;
  const int waves = (block + 63) / 64;
  const int by_lds = lds > 0 ? (int)((160 * 1024) / lds) : 32;
  const int own = std::max(1, std::min(by_lds, 32 / std::max(waves, 1)));
  return std::max(per_cu, own) * std::max(cus, 1);
}

// The slab and family-backprojection launches of the items of NB-image
// slots (k = 0: NB 8, the table's first n8 items; k = 1: NB 4, the rest).
template <int F, int NB, int R>
int launch_step_nb(fir_osem_step* st, int k, const int2* items, int n)
{
  const StepJob* jobs = reinterpret_cast<const StepJob*>(st->d_jobs);
  auto kern = k_step_slab<F, NB, R, 2>;
  static size_t allowed = 0;
  if (st->slab_lds > 65536 && allowed < st->slab_lds &&
      hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                          hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)st->slab_lds) == hipSuccess)
    allowed = st->slab_lds;
  if (st->grid_slab[k] == 0)
  {
    const int res = resident_blocks(reinterpret_cast<const void*>(kern),
                                    st->slab_block, st->slab_lds,
                                    st->ctx->device);
    st->grid_slab[k] = std::max(1, std::min(n, res > 0 ? res : 512));
  }
  hipLaunchKernelGGL(kern, dim3(st->grid_slab[k]), dim3(st->slab_block),
                     st->slab_lds, st->ctx->stream, jobs, items, n,
                     st->d_ctr + 4 * k);
  FIR_CHECK_LAUNCH();
  auto bp = k_step_fam_bp<F, NB>;
  if (st->grid_bp[k] == 0)
  {
    const int res = resident_blocks(reinterpret_cast<const void*>(bp),
                                    kSheetBlock,
                                    sizeof(float) * kFamRowsMax * NB,
                                    st->ctx->device);
    st->grid_bp[k] = std::max(1, std::min(n, res > 0 ? res : 512));
  }
  hipLaunchKernelGGL(bp, dim3(st->grid_bp[k]), dim3(kSheetBlock), 0,
                     st->ctx->stream, jobs, items, n, st->d_ctr + 4 * k + 2);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

template <int F, int NB>
int launch_step_r(fir_osem_step* st, int k, const int2* items, int n)
{
  switch (st->slab_R)
  {
  case 1: return launch_step_nb<F, NB, 1>(st, k, items, n);
  case 2: return launch_step_nb<F, NB, 2>(st, k, items, n);
  case 3: return launch_step_nb<F, NB, 3>(st, k, items, n);
  default: return launch_step_nb<F, NB, 5>(st, k, items, n);
  }
}

template <int F>
int launch_step_projector(fir_osem_step* st)
{
  hipStream_t s = st->ctx->stream;
  const StepJob* jobs = reinterpret_cast<const StepJob*>(st->d_jobs);
  hipLaunchKernelGGL(k_step_pack, dim3(st->pack_blocks), dim3(256), 0, s,
                     jobs, st->d_pack_pre, st->n_jobs);
  FIR_CHECK_LAUNCH();
  int rc;
  const int n8 = st->n_items8, n4 = st->n_items - st->n_items8;
  if (n8 > 0 && (rc = launch_step_r<F, 8>(st, 0, st->d_items, n8)))
    return rc;
  if (n4 > 0 && (rc = launch_step_r<F, 4>(st, 1, st->d_items + n8, n4)))
    return rc;
  hipLaunchKernelGGL(k_step_reduce, dim3(st->red_blocks), dim3(kRedBlock),
                     st->red_tile, s, jobs, st->d_red_pre, st->n_jobs,
                     st->red_tile > 0 ? 1 : 0);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

} // namespace

extern "C" {

int fir_cache_job_workspace(const fir_cache* cache, int64_t* bytes)
{
  int rc = check_cache(cache);
  if (rc) return rc;
  if (!bytes) return fail(FIR_E_ARG, "NULL output");
  *bytes = (int64_t)ws_layout(cache).total;
  return FIR_OK;
}

int fir_osem_step_create(int n_jobs, const fir_osem_job* jobs,
                         const float* fwhm, float cut, fir_osem_step** out)
{
  if (!out) return fail(FIR_E_ARG, "NULL output");
  *out = nullptr;
  if (n_jobs <= 0 || !jobs) return fail(FIR_E_ARG, "no jobs");
  int rc;
  fir_ctx* c = nullptr;
  int fmt = -1, lanes = 0, max_pairs = 0;
  for (int j = 0; j < n_jobs; ++j)
  {
    const fir_osem_job& J = jobs[j];
    if ((rc = check_cache(J.cache))) return rc;
    const fir_cache* cache = J.cache;
    const RowMatrix* rm = cache->rm.get();
    if (c && cache->ctx != c)
      return fail(FIR_E_ARG, "job %d: the caches of a step must share one "
                             "context", j);
    c = cache->ctx;
    if (J.subset < 0 || J.subset >= cache->n_cached)
      return fail(FIR_E_ARG, "job %d: subset %d is not held by its cache", j,
                  J.subset);
    if (J.n_images < 1 || J.n_images > kMaxCacheBatch)
      return fail(FIR_E_ARG, "job %d: n_images %d outside 1..%d", j,
                  J.n_images, kMaxCacheBatch);
    for (int b = 0; b < J.n_images; ++b)
      if (!J.x[b] || !J.y[b] || !J.bp[b])
        return fail(FIR_E_ARG, "job %d: NULL image, sinogram or scratch", j);
    if (!J.sens) return fail(FIR_E_ARG, "job %d: NULL sensitivity", j);
    if (!J.workspace ||
        (reinterpret_cast<uintptr_t>(J.workspace) & (kWsAlign - 1)))
      return fail(FIR_E_ARG, "job %d: workspace NULL or not 256-byte aligned",
                  j);
    for (int k = 0; k < j; ++k)
      if (jobs[k].workspace == J.workspace)
        return fail(FIR_E_ARG, "jobs %d and %d share a workspace", k, j);
    if (!use_slabs(cache, kMaxCacheBatch) || cache->h_ford.empty())
      return fail(FIR_E_ARG, "job %d: its cache lacks the sheet form, family "
                             "lists or slab form (fir_cache_forms)", j);
    if (fmt >= 0 && rm->fmt != fmt)
      return fail(FIR_E_ARG, "job %d: the caches' length formats differ", j);
    fmt = rm->fmt;
    if (lanes && cache->fwd_lanes != lanes)
      return fail(FIR_E_ARG, "job %d: the caches' lanes per LOR differ", j);
    lanes = cache->fwd_lanes;
    max_pairs = std::max(max_pairs, rm->n_pairs);
  }
  const bool gauss =
    fwhm && fwhm[0] > 0.0f && fwhm[1] > 0.0f && fwhm[2] > 0.0f;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  std::unique_ptr<fir_osem_step> st(new fir_osem_step);
  st->ctx = c;
  st->fmt = fmt;
  st->n_jobs = n_jobs;
  st->gauss = gauss;
  st->cut = cut;
  if (gauss)
  {
    // as fir_convolve + fir_cut_circle: the edge restore beyond
    // cut - 5 * max(FWHM x, y), then the cut (operations.cc:158-220)
    const float fw = fwhm[0] > fwhm[1] ? fwhm[0] : fwhm[1];
    st->restore = cut > 0.0f ? 1 : 0;
    st->restore_r = (double)cut - 5.0 * fw;
  }
  // slab launch shape: one block size for every job (the same rows per
  // family cover every job's families)
  {
    SlabShape best;
    for (int R : kSlabRounds)
    {
      const int per_round = (max_pairs + R - 1) / R;
      const int block = (per_round * lanes + 63) / 64 * 64;
      if (block <= kSlabBlockMax)
      {
        best.R = R;
        best.block = block;
        break;
      }
    }
    if (best.R == 0)
      return fail(FIR_E_ARG, "families of %d rows exceed the slab kernel",
                  max_pairs);
    st->slab_R = best.R;
    st->slab_block = best.block;
  }
  std::vector<StepJob> hj(n_jobs);
  std::vector<std::pair<long long, int2>> items; // (cost, (job, family))
  std::vector<int> pack_pre(n_jobs + 1), red_pre(n_jobs + 1);
  std::vector<firi::StepImage> imgs;
  int pack = 0, red = 0, tiles = 0;
  size_t red_tile = 0;
  bool red_fit = true;
  int vtiles[3] = {0, 0, 0};
  bool vec = true;
  for (int j = 0; j < n_jobs; ++j)
  {
    const fir_osem_job& J = jobs[j];
    fir_cache* cache = J.cache;
    const RowMatrix* rm = cache->rm.get();
    const fir_vol_desc& v = cache->vol;
    const int nxy = v.nx * v.ny, nz = v.nz;
    const WsLayout L = ws_layout(cache);
    char* ws = static_cast<char*>(J.workspace);
    const int nb = J.n_images <= 4 ? 4 : 8;
    const SlabShape sh = slab_shape(cache, nb);
    StepJob& s = hj[j];
    std::memset(&s, 0, sizeof(s));
    s.fam_row = rm->d_fam_row;
    s.fam_off = rm->d_fam_off;
    s.fam_xy = rm->d_fam_xy;
    s.sel = rm->d_sel;
    s.rd = row_data(rm);
    s.fl_slot = rm->d_fl_slot;
    s.fl_gb = rm->d_fl_gb;
    s.fl = rm->d_fl;
    s.fl_nact = rm->d_fl_nact;
    s.fam_loc = cache->d_fam_loc;
    s.fr_off = cache->d_fr_off + (long long)J.subset * (nxy + 1);
    s.fr_ent = cache->d_fr_ent;
    s.w = reinterpret_cast<float*>(ws + L.w);
    s.X = reinterpret_cast<float*>(ws + L.x);
    s.part = reinterpret_cast<float*>(ws + L.part);
    s.io.n_images = J.n_images;
    for (int b = 0; b < J.n_images; ++b)
    {
      s.io.y[b] = J.y[b];
      s.io.bias[b] = J.bias[b];
      s.in.p[b] = J.x[b];
      s.out.p[b] = J.bp[b];
    }
    s.nxy = nxy;
    s.nz = nz;
    s.lanes = lanes;
    s.slab_px = sh.slab_px;
    s.buf_f4 = sh.buf_f4;
    s.pack_tiles_x = (nxy + kPzPx - 1) / kPzPx;
    s.nb = nb;
    st->slab_lds = std::max(st->slab_lds, sh.lds);
    {
      // the reduction's output tile: staged only if every job's fits
      const size_t tb = red_tile_bytes(nb, nz);
      red_fit = red_fit && tb > 0;
      red_tile = std::max(red_tile, tb);
    }
    pack_pre[j] = pack;
    pack += s.pack_tiles_x * ((nz + kPzZ - 1) / kPzZ);
    red_pre[j] = red;
    red += (nxy + kRedBlock / 64 - 1) / (kRedBlock / 64);
    const int f0 = cache->ford_first[J.subset];
    const int f1 = cache->ford_first[J.subset + 1];
    for (int k = f0; k < f1; ++k)
    {
      const int f = cache->h_ford[k];
      items.push_back({cache->h_fcost[f], make_int2(j, f)});
    }
    const float* taps = nullptr;
    int h[3] = {0, 0, 0};
    if (gauss && (rc = firi::gauss_taps_for(c, &v, fwhm, &taps, h)))
      return rc;
    for (int b = 0; b < J.n_images; ++b)
    {
      firi::StepImage m;
      std::memset(&m, 0, sizeof(m));
      m.x = J.x[b];
      m.bp = J.bp[b];
      m.sens = J.sens;
      m.t1 = reinterpret_cast<float*>(ws + L.t1) + (size_t)b * cache->nvox;
      m.t2 = reinterpret_cast<float*>(ws + L.t2) + (size_t)b * cache->nvox;
      m.taps = taps;
      for (int d = 0; d < 3; ++d) m.h[d] = h[d];
      m.nx = v.nx;
      m.ny = v.ny;
      m.nz = v.nz;
      m.dx = v.dx;
      m.dy = v.dy;
      m.tile0 = tiles;
      tiles += firi::step_image_tiles(v.nx, v.ny, v.nz);
      for (int d = 0; d < 3; ++d)
      {
        m.vt0[d] = vtiles[d];
        vtiles[d] += firi::step_image_vtiles(v.nx, v.ny, v.nz, d);
      }
      // the vectorised passes: rows of whole float4s, 16-byte aligned
      // images, half-widths they handle and that every image shares
      const uintptr_t al = reinterpret_cast<uintptr_t>(m.x) |
        reinterpret_cast<uintptr_t>(m.t1) | reinterpret_cast<uintptr_t>(m.t2);
      if (!gauss || !firi::gauss_vec_shape(v.nx, h) || (al & 15u) != 0 ||
          (!imgs.empty() && (imgs[0].h[0] != h[0] || imgs[0].h[1] != h[1] ||
                             imgs[0].h[2] != h[2])))
        vec = false;
      imgs.push_back(m);
    }
  }
  pack_pre[n_jobs] = pack;
  red_pre[n_jobs] = red;
  // the jobs of 8-image slots first, each class the most work first; ties
  // in job then family order
  std::stable_sort(items.begin(), items.end(),
                   [&](const std::pair<long long, int2>& a,
                       const std::pair<long long, int2>& b) {
                     const int ka = hj[a.second.x].nb == 4;
                     const int kb = hj[b.second.x].nb == 4;
                     return ka != kb ? ka < kb : a.first > b.first;
                   });
  for (const auto& x : items)
    if (hj[x.second.x].nb == 8) ++st->n_items8;
  std::vector<int2> it(items.size());
  for (size_t k = 0; k < items.size(); ++k) it[k] = items[k].second;
  st->n_items = (int)it.size();
  st->pack_blocks = pack;
  st->red_blocks = red;
  st->red_tile = red_fit ? red_tile : 0;
  st->n_imgs = (int)imgs.size();
  st->img_tiles = tiles;
  st->vec_gauss = vec && !imgs.empty();
  for (int d = 0; d < 3; ++d)
  {
    st->vtiles[d] = vtiles[d];
    st->vh[d] = imgs.empty() ? 0 : imgs[0].h[d];
  }
  StepJob* d_jobs = nullptr;
  if ((rc = dev_alloc(&d_jobs, (size_t)n_jobs, "step jobs")))
    return rc;
  st->d_jobs = d_jobs;
  if (
      (rc = dev_alloc(&st->d_items, std::max(it.size(), (size_t)1),
                      "step items")) ||
      (rc = dev_alloc(&st->d_pack_pre, (size_t)n_jobs + 1, "step pack")) ||
      (rc = dev_alloc(&st->d_red_pre, (size_t)n_jobs + 1, "step reduce")) ||
      (rc = dev_alloc(&st->d_imgs, imgs.size(), "step images")) ||
      (rc = dev_alloc(&st->d_ctr, (size_t)8, "step queues")))
    return rc;
  hipStream_t s = c->stream;
  FIR_HIP_CHECK(hipMemsetAsync(st->d_ctr, 0, 8 * sizeof(int), s));
  FIR_HIP_CHECK(hipMemcpyAsync(st->d_jobs, hj.data(), sizeof(StepJob) * n_jobs,
                               hipMemcpyHostToDevice, s));
  if (!it.empty())
    FIR_HIP_CHECK(hipMemcpyAsync(st->d_items, it.data(),
                                 sizeof(int2) * it.size(),
                                 hipMemcpyHostToDevice, s));
  FIR_HIP_CHECK(hipMemcpyAsync(st->d_pack_pre, pack_pre.data(),
                               sizeof(int) * (n_jobs + 1),
                               hipMemcpyHostToDevice, s));
  FIR_HIP_CHECK(hipMemcpyAsync(st->d_red_pre, red_pre.data(),
                               sizeof(int) * (n_jobs + 1),
                               hipMemcpyHostToDevice, s));
  FIR_HIP_CHECK(hipMemcpyAsync(st->d_imgs, imgs.data(),
                               sizeof(firi::StepImage) * imgs.size(),
                               hipMemcpyHostToDevice, s));
  FIR_HIP_CHECK(hipStreamSynchronize(s));
  *out = st.release();
  return FIR_OK;
}

int fir_osem_step_run(fir_osem_step* st)
{
  if (!st || !st->ctx) return fail(FIR_E_ARG, "step is NULL");
  FIR_HIP_CHECK(hipSetDevice(st->ctx->device));
  const int rc = st->fmt == FIR_LEN_SNORM16
    ? launch_step_projector<FIR_LEN_SNORM16>(st)
    : launch_step_projector<FIR_LEN_FP16>(st);
  if (rc) return rc;
  return firi::launch_step_elementwise(
    st->ctx->stream, st->d_imgs, st->n_imgs, st->img_tiles, st->gauss,
    st->cut, st->restore, st->restore_r, st->vec_gauss ? st->vtiles : nullptr,
    st->vh);
}

int fir_osem_step_destroy(fir_osem_step* st)
{
  delete st;
  return FIR_OK;
}

} // extern "C"
