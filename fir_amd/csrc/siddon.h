// siddon.h -- device-side LOR geometry and Siddon walk for gfx950.
//
// The walk reproduces the reference's FP64 parametric Siddon walk
// (Siddon.cc:112-412) operation for operation, so that the voxel-index
// sequence of every LOR is bit-identical to the reference's and the lengths
// are the same doubles:
//   - per-dimension setup with the |diff| <= 1e-5 "cancelled dimension"
//     rule (Siddon.cc:260-317),
//   - entry/exit alphas as max/min over dimensions and [0, 1]
//     (Siddon.cc:319-334),
//   - start voxel by truncation and clamp (Siddon.cc:336-350),
//   - first crossing per dimension (Siddon.cc:352-377),
//   - the loop: next = min(alphaMax, ax, ay, az); emit when in bounds;
//     advance every dimension whose crossing is within 1e-5 of next
//     (Siddon.cc:205-252, 379-392).
// The translation unit is compiled with -ffp-contract=off: every multiply
// and add rounds separately, as on the reference's x86-64 build. FP64
// division and sqrt on gfx950 are correctly rounded, like x86 divsd/sqrtsd
// (checked on 16.7 M random operands: no difference).
#pragma once

#include <hip/hip_runtime.h>

namespace fir
{

constexpr double kEps = 0.00001;
// (double)x > 1e-5 holds exactly when the float x > 1e-5f: float(1e-5) is
// the float just below 1e-5 and the next float is above it.
constexpr float kEpsF = 0.00001f;

// Walk constants of one volume (reference Siddon ctor, Siddon.cc:20-52).
struct Grid
{
  double low[3];
  double high[3];
  double step[3]; // int-truncated extents (reference) or true extents
  int nm1[3];
  int row;        // nx
  int slice;      // nx * ny
  int max_iter;   // nx + ny + nz + 8: bound on the walk loop
};

// Sinogram / scanner geometry needed to turn a work index into a LOR.
// The (segment, axial) rows are visited in the order of `pairs`; within a
// pair the views of the subset, within a view the tangential bins.
struct LorGeom
{
  const double2* xy;         // [n_cpr] crystal (x, y)
  const double* z;           // [n_slices] slice z
  const int4* pairs;         // [n_pairs] (segment rank, axial, slice1, slice2)
  const int* n_axial;        // [n_seg] axial count per segment rank
  const long long* seg_base; // [n_seg] first bin of each segment
  int n_views;
  int n_tang;
  int tang_off;
  int n_cpr;
  int n_pairs;
  const int* ax_prefix; // [n_seg + 1] first pair of each segment (segment-ordered pairs)
  int n_seg;
  int row_mode;   // ROWS_* below
  int tang_major; // lanes: consecutive rows at one tangential bin (1) or consecutive bins of one row (0)
};

// Row orders of a subset's (view, pair) rows.
enum
{
  ROWS_PAIR_MAJOR = 0, // all views of one pair together
  ROWS_VIEW_MAJOR = 1, // all pairs of one view together
  ROWS_SEGMENT_VIEW = 2 // segment, then view, then axial (pairs in segment order)
};

struct Lor
{
  double e1[3];
  double e2[3];
  long long bin;
};

// Views of one subset: v = subview * n_subsets + subset, v < n_views
// (reference LORCache.cc:49-51). With a divisor n_subsets this is
// n_views / n_subsets for every subset; otherwise the first subsets get one
// view more.
__host__ __device__ __forceinline__ int views_in_subset(int n_views,
                                                        int n_subsets,
                                                        int subset)
{
  return n_subsets > 0 ? (n_views - subset + n_subsets - 1) / n_subsets
                       : n_views;
}

// (view, tangential bin) -> the LOR's two angular crystals, the walk
// running from the first to the second (ProjData.cc:301-340); C++ integer
// division truncates toward zero.
__device__ __forceinline__ int2 view_crystals(const LorGeom& g, int view,
                                              int tt)
{
  const int tang = tt - g.tang_off;
  const int shift = tang >= 0 ? tang / 2 : (tang - 1) / 2;
  const int odd = (tang < 0 ? -tang : tang) & 1;
  int c1 = view + shift + odd;
  int c2 = view + g.n_views - shift;
  if (c1 < 0)
    c1 += g.n_cpr;
  else if (c1 >= g.n_cpr)
    c1 -= g.n_cpr;
  if (c2 < 0)
    c2 += g.n_cpr;
  else if (c2 >= g.n_cpr)
    c2 -= g.n_cpr;
  return make_int2(c1, c2);
}

// Work index t -> LOR (t < 2^31, checked by the host). n_subsets == 0
// selects every view. The bin index is the LOR's position in the
// reference's segment-concatenated sinogram (ProjData.h:9-21).
__device__ __forceinline__ void lor_from_index(const LorGeom& g, long long t,
                                               int n_subsets, int subset,
                                               Lor& L)
{
  const unsigned vpr = (unsigned)views_in_subset(g.n_views, n_subsets, subset);
  const unsigned ut = (unsigned)t;
  unsigned row;
  int tt;
  if (g.tang_major)
  {
    const unsigned n_rows = vpr * (unsigned)g.n_pairs;
    const unsigned q = ut / n_rows;
    tt = (int)q;
    row = ut - q * n_rows;
  }
  else
  {
    row = ut / (unsigned)g.n_tang;
    tt = (int)(ut - row * (unsigned)g.n_tang);
  }
  unsigned p, svu;
  if (g.row_mode == ROWS_SEGMENT_VIEW)
  {
    // Segment k holds rows [vpr * ax_prefix[k], vpr * ax_prefix[k + 1]).
    int lo = 0, hi = g.n_seg - 1;
    while (lo < hi)
    {
      const int mid = (lo + hi + 1) >> 1;
      if (vpr * (unsigned)g.ax_prefix[mid] <= row)
        lo = mid;
      else
        hi = mid - 1;
    }
    const unsigned first = (unsigned)g.ax_prefix[lo];
    const unsigned na = (unsigned)g.ax_prefix[lo + 1] - first;
    const unsigned local = row - vpr * first;
    svu = local / na;
    p = first + (local - svu * na);
  }
  else if (g.row_mode == ROWS_VIEW_MAJOR)
  {
    svu = row / (unsigned)g.n_pairs;
    p = row - svu * (unsigned)g.n_pairs;
  }
  else
  {
    p = row / vpr;
    svu = row - p * vpr;
  }
  const int sv = (int)svu;
  const int4 P = g.pairs[p];
  const int na = g.n_axial[P.x];
  const int view = n_subsets > 0 ? sv * n_subsets + subset : sv;
  L.bin = g.seg_base[P.x] + ((long long)view * na + P.y) * g.n_tang + tt;

  const int2 c = view_crystals(g, view, tt);
  const double2 p1 = g.xy[c.x];
  const double2 p2 = g.xy[c.y];
  L.e1[0] = p1.x;
  L.e1[1] = p1.y;
  L.e1[2] = g.z[P.z];
  L.e2[0] = p2.x;
  L.e2[1] = p2.y;
  L.e2[2] = g.z[P.w];
}

// Bin index (sinogram order) -> LOR, for parity dumps: the segment by
// binary search of seg_base, then view, axial and tangential bin. `pairs`
// must be in sinogram (segment, axial) order and ax_prefix its per-segment
// offsets.
__device__ __forceinline__ void lor_from_bin(const LorGeom& g,
                                             const int* ax_prefix, int n_seg,
                                             long long b, Lor& L)
{
  int lo = 0, hi = n_seg;
  while (hi - lo > 1)
  {
    const int mid = (lo + hi) >> 1;
    if (g.seg_base[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  const int k = lo;
  const int na = g.n_axial[k];
  const long long r = b - g.seg_base[k];
  const long long per_view = (long long)na * g.n_tang;
  const int view = (int)(r / per_view);
  const long long r2 = r - (long long)view * per_view;
  const int axial = (int)(r2 / g.n_tang);
  const int tt = (int)(r2 - (long long)axial * g.n_tang);
  const int4 P = g.pairs[ax_prefix[k] + axial];
  L.bin = b;
  const int2 c = view_crystals(g, view, tt);
  const double2 p1 = g.xy[c.x];
  const double2 p2 = g.xy[c.y];
  L.e1[0] = p1.x;
  L.e1[1] = p1.y;
  L.e1[2] = g.z[P.z];
  L.e2[0] = p2.x;
  L.e2[1] = p2.y;
  L.e2[2] = g.z[P.w];
}

// Target of a scatter: the image itself, or one of `copies` private
// copies (stride apart) chosen by the XCD the wave runs on, so that each
// XCD's L2 holds its own copy's lines. The copies are summed afterwards.
struct Scatter
{
  float* base;
  long long stride;
  int copies;
};

__device__ __forceinline__ float* scatter_target(const Scatter& s)
{
  if (s.copies <= 1) return s.base;
  const unsigned xcc = __builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11));
  return s.base + (long long)(xcc % (unsigned)s.copies) * s.stride;
}

// Walk state after set-up; copying it replays the identical walk.
struct Walk
{
  double a[3];
  double da[3];
  double prev;
  double a_hi;
  double d12;
  int pos[3];
  int dir[3];
};

__device__ __forceinline__ double dmin(double a, double b)
{
  return a < b ? a : b;
}

__device__ __forceinline__ double dmax(double a, double b)
{
  return a > b ? a : b;
}

__device__ __forceinline__ double dabs(double a) { return a < 0 ? -a : a; }

// Set-up of the walk between e1 and e2. Returns false if the line misses the
// volume (the reference's computePath returns false).
__device__ __forceinline__ bool walk_setup(const Grid& g, const double* c1,
                                           const double* c2, Walk& w)
{
  double diff[3], amin[3], amax[3], nxt[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    double df = c2[d] - c1[d];
    if (dabs(df) > kEps)
    {
      if (df > 0)
      {
        w.dir[d] = 1;
        amin[d] = (g.low[d] - c1[d]) / df;
        amax[d] = (g.high[d] - c1[d]) / df;
      }
      else
      {
        w.dir[d] = -1;
        amin[d] = (g.high[d] - c1[d]) / df;
        amax[d] = (g.low[d] - c1[d]) / df;
      }
      nxt[d] = 0.0;
    }
    else
    {
      if (c1[d] < g.low[d] || c1[d] > g.high[d]) return false;
      df = kEps;
      w.dir[d] = -1;
      amin[d] = 0.0;
      amax[d] = 1.0;
      nxt[d] = 1.0;
    }
    diff[d] = df;
  }
  const double a_lo = dmax(amin[0], dmax(amin[1], dmax(amin[2], 0.0)));
  const double a_hi = dmin(amax[0], dmin(amax[1], dmin(amax[2], 1.0)));
  if (a_lo >= a_hi) return false;
#pragma unroll
  for (int d = 0; d < 3; ++d) w.da[d] = g.step[d] / dabs(diff[d]);
  w.d12 = sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    int i = (int)((c1[d] + diff[d] * a_lo - g.low[d]) / g.step[d]);
    i = i < g.nm1[d] ? i : g.nm1[d];
    i = 0 > i ? 0 : i;
    w.pos[d] = i;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    double al = nxt[d];
    if (al < 1.0) al = (g.low[d] + g.step[d] * w.pos[d] - c1[d]) / diff[d];
    if (w.dir[d] > 0) al += w.da[d];
    w.a[d] = al;
  }
  w.prev = a_lo;
  w.a_hi = a_hi;
  return true;
}

// Run the walk, calling emit(linear_index, length) for every in-volume
// element in path order.
template <class Emit>
__device__ __forceinline__ void walk_run(const Grid& g, Walk w, Emit&& emit)
{
  double prev = w.prev;
  for (int it = 0; prev < w.a_hi && it < g.max_iter; ++it)
  {
    const double next = dmin(w.a_hi, dmin(w.a[0], dmin(w.a[1], w.a[2])));
    if ((unsigned)w.pos[0] <= (unsigned)g.nm1[0] &&
        (unsigned)w.pos[1] <= (unsigned)g.nm1[1] &&
        (unsigned)w.pos[2] <= (unsigned)g.nm1[2])
    {
      emit(w.pos[0] + w.pos[1] * g.row + w.pos[2] * g.slice,
           (next - prev) * w.d12);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      if (dabs(w.a[d] - next) < kEps)
      {
        w.a[d] += w.da[d];
        w.pos[d] += w.dir[d];
      }
    }
    prev = next;
  }
}

// Gather walk, software-pipelined by one element: at step k the voxel
// load of element k is issued (fetch) before element k-1 is accumulated
// (consume), so each load's latency overlaps one step of walk arithmetic.
// Elements are consumed in path order, so accumulations are unchanged.
// fetch(idx) must accept idx = 0 (used for out-of-volume steps, whose value
// is ignored); consume(value, length, valid) must ignore calls with
// valid == false.
template <class Fetch, class Consume>
__device__ __forceinline__ void walk_gather(const Grid& g, Walk w,
                                            Fetch&& fetch, Consume&& consume)
{
  double prev = w.prev;
  bool have = false;
  auto pv = fetch(0);
  double plen = 0.0;
  for (int it = 0; prev < w.a_hi && it < g.max_iter; ++it)
  {
    const double next = dmin(w.a_hi, dmin(w.a[0], dmin(w.a[1], w.a[2])));
    const bool inb = (unsigned)w.pos[0] <= (unsigned)g.nm1[0] &&
      (unsigned)w.pos[1] <= (unsigned)g.nm1[1] &&
      (unsigned)w.pos[2] <= (unsigned)g.nm1[2];
    const int idx = inb ? w.pos[0] + w.pos[1] * g.row + w.pos[2] * g.slice : 0;
    const auto v = fetch(idx);
    const double len = (next - prev) * w.d12;
    consume(pv, plen, have);
    pv = v;
    plen = len;
    have = inb;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      if (dabs(w.a[d] - next) < kEps)
      {
        w.a[d] += w.da[d];
        w.pos[d] += w.dir[d];
      }
    }
    prev = next;
  }
  consume(pv, plen, have);
}

} // namespace fir
