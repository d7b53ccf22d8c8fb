// listmode.hip -- the sinogram's wire formats on the device: list-mode
// events (crystal pairs) histogrammed into sinogram bins, and the
// span > 1 weighting of fused ring pairs.
//
// The reference does both on the host, one call per event or bin:
// ProjData::getBinCoordinates (ProjData.cc:342-479) + incrementBin
// (ProjData.inl:47-56), and ProjData::rebinWeight (ProjData.cc:541-615).
// Here an event is one lane: its bin comes from integer arithmetic on the
// sinogram geometry, and the count is an FP32 atomic add (exact for counts
// below 2^24 per bin). The rebin weight of a bin is a closed form of its
// (segment, axial) position: the number of ring pairs the segment fuses
// there.
#include "internal.h"

using namespace firi;

namespace
{

// Sinogram geometry as the kernels need it.
struct BinGeom
{
  int n_rings, n_cpr, span, half_span;
  int n_seg, seg_off, n_tang, tang_off, n_views, max_ring_diff;
  long long n_bins;
  const int* n_axial;       // [n_seg]
  const long long* seg_base; // [n_seg] first bin of each segment
};

// Largest ring difference the stored segments hold (ProjHeader.cc:169-186).
int max_ring_diff(const fir_ctx* c)
{
  if (c->span == 1) return c->seg_off;
  return c->half_span + (c->seg_off > 0 ? c->seg_off * c->span : 0);
}

// Crystals per ring from the sinogram geometry itself (nViews =
// nCrystalsPerRing / 2, ProjHeader.cc:155-156), so that no scanner tables
// are needed.
BinGeom bin_geom(const fir_ctx* c)
{
  return BinGeom{c->n_rings, 2 * c->n_views, c->span, c->half_span,
                 c->n_seg,   c->seg_off,  c->n_tang,   c->tang_off,
                 c->n_views, max_ring_diff(c), (long long)c->n_bins,
                 c->d_n_axial, c->d_seg_base};
}

// First ring sum of fused segment |seg| = m > 0 (span > 1): the axial
// coordinate of a ring pair is its ring sum minus this.
__device__ __forceinline__ int first_ring_sum(const BinGeom& g, int m)
{
  return m == 0 ? 0 : 1 + g.half_span + (m - 1) * g.span;
}

// Flat bin of the crystal pair (ring r1, crystal c1) - (ring r2, crystal
// c2), or -1 if the sinogram does not hold it (getBinCoordinates,
// ProjData.cc:342-479; the pair's order does not matter).
__device__ long long bin_of_pair(const BinGeom& g, int r1, int c1, int r2,
                                 int c2)
{
  const int n = g.n_cpr;
  const int d = r1 > r2 ? r1 - r2 : r2 - r1;
  if (d > g.max_ring_diff) return -1;
  // Tangential coordinate: signed distance from n/2 of the crystal
  // separation, the sign given by the half-turn the crystal sum lies in.
  const int sum = c1 + c2;
  const int s1 = (sum >= n / 2 && sum < 3 * n / 2) ? -1 : 1;
  const int t = s1 * ((c2 > c1 ? c2 - c1 : c1 - c2) - n / 2);
  if (t < -g.tang_off || t >= g.n_tang - g.tang_off) return -1;
  int m, axial;
  if (g.span == 1)
  {
    m = d;
    axial = (r1 + r2 - d) / 2;
  }
  else
  {
    m = d <= g.half_span ? 0 : 1 + (d - g.half_span - 1) / g.span;
    axial = r1 + r2 - first_ring_sum(g, m);
  }
  const int view = ((sum + n / 2) % n) / 2;
  int seg = m;
  if (m != 0)
  {
    // Sign: move both crystals to the parallel LOR through the centre
    // (tangential 0 or 1) and compare them, and the rings.
    const int s2 = c1 < c2 ? 1 : -1;
    const int u = (t % 2 == 0) ? -t / 2 : -(t - 1) / 2;
    int k1 = c1 - s1 * s2 * u;
    int k2 = c2 + s1 * s2 * u;
    k1 += k1 >= n ? -n : (k1 < 0 ? n : 0);
    k2 += k2 >= n ? -n : (k2 < 0 ? n : 0);
    seg = m * (r1 < r2 ? -1 : 1) * (k1 < k2 ? 1 : -1);
  }
  const int k = seg + g.seg_off;
  if (k < 0 || k >= g.n_seg) return -1;
  const int na = g.n_axial[k];
  if (axial < 0 || axial >= na || view < 0 || view >= g.n_views) return -1;
  return g.seg_base[k] + ((long long)view * na + axial) * g.n_tang + t +
    g.tang_off;
}

// sino[bin] += w (or 1) per event; per-wave counts of the events kept go
// to *counted.
__global__ void k_histogram(BinGeom g, const int4* __restrict__ ev,
                            const float* __restrict__ w, long long n,
                            float* __restrict__ sino,
                            unsigned long long* __restrict__ counted)
{
  for (long long i0 = (long long)blockIdx.x * blockDim.x;; i0 +=
       (long long)gridDim.x * blockDim.x)
  {
    // every lane of a block takes part in each round, so the wave-level
    // count below sees whole waves
    if (i0 >= n) break;
    const long long i = i0 + threadIdx.x;
    bool kept = false;
    if (i < n)
    {
      const int4 e = ev[i];
      if (e.x >= 0 && e.x < g.n_rings && e.z >= 0 && e.z < g.n_rings &&
          e.y >= 0 && e.y < g.n_cpr && e.w >= 0 && e.w < g.n_cpr)
      {
        const long long b = bin_of_pair(g, e.x, e.y, e.z, e.w);
        if (b >= 0)
        {
          atomicAdd(sino + b, w ? w[i] : 1.0f);
          kept = true;
        }
      }
    }
    if (counted)
    {
      const unsigned long long mask = __ballot(kept);
      if ((threadIdx.x & (warpSize - 1)) == 0 && mask)
        atomicAdd(counted, (unsigned long long)__popcll(mask));
    }
  }
}

// Integers of [a, b] with the parity of p.
__device__ __forceinline__ int count_parity(int a, int b, int p)
{
  const int first = a + (((a - p) % 2 + 2) % 2);
  return first > b ? 0 : (b - first) / 2 + 1;
}

// Ring pairs (ordered as the segment's sign orders them) that segment seg
// fuses at axial position ax: ring sums are fixed by (seg, ax); ring
// differences run over the segment's range, within the rings and the
// stored segments, with the parity of the ring sum.
__device__ int fused_pairs(const BinGeom& g, int seg, int ax)
{
  const int m = seg < 0 ? -seg : seg;
  const int rs = ax + first_ring_sum(g, m);
  int lim = rs < 2 * (g.n_rings - 1) - rs ? rs : 2 * (g.n_rings - 1) - rs;
  lim = lim < g.max_ring_diff ? lim : g.max_ring_diff;
  if (lim < 0) return 0;
  if (m == 0)
  {
    const int h = g.half_span < lim ? g.half_span : lim;
    return count_parity(-h, h, rs);
  }
  const int lo = g.half_span + 1 + (m - 1) * g.span;
  int hi = g.half_span + m * g.span;
  hi = hi < lim ? hi : lim;
  return count_parity(lo, hi, rs);
}

// sino[b] *= (float)(1.0 / k) for every bin of a (segment, axial) position
// that fuses k > 1 ring pairs (span > 1).
__global__ void k_rebin_weight(BinGeom g, float* __restrict__ sino)
{
  for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
       b < g.n_bins; b += (long long)gridDim.x * blockDim.x)
  {
    int lo = 0, hi = g.n_seg - 1; // last segment starting at or before b
    while (lo < hi)
    {
      const int mid = (lo + hi + 1) / 2;
      if (g.seg_base[mid] <= b) lo = mid;
      else hi = mid - 1;
    }
    const long long r = (b - g.seg_base[lo]) / g.n_tang;
    const int ax = (int)(r % g.n_axial[lo]);
    const int k = fused_pairs(g, lo - g.seg_off, ax);
    if (k > 1) sino[b] *= (float)(1.0 / (double)k);
  }
}

} // namespace

extern "C" {

int fir_histogram_events(fir_ctx* c, const int* events, int64_t n_events,
                         const float* weights, float* sino,
                         int64_t* n_counted)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (!c->proj_set) return fail(FIR_E_STATE, "sinogram geometry not set");
  if (n_events < 0) return fail(FIR_E_ARG, "negative event count");
  if (n_events > 0 && (!events || !sino))
    return fail(FIR_E_ARG, "NULL buffer");
  if (((uintptr_t)events & 15) != 0)
    return fail(FIR_E_ARG, "events must be 16-byte aligned (4 int32 each)");
  FIR_HIP_CHECK(hipSetDevice(c->device));
  if (n_counted)
  {
    if (!c->d_counter)
      FIR_HIP_CHECK(hipMalloc(&c->d_counter, sizeof(unsigned long long)));
    FIR_HIP_CHECK(hipMemsetAsync(c->d_counter, 0, sizeof(unsigned long long),
                                 c->stream));
  }
  if (n_events > 0)
  {
    hipLaunchKernelGGL(k_histogram, dim3(ew_blocks(n_events)), dim3(kBlock),
                       0, c->stream, bin_geom(c),
                       reinterpret_cast<const int4*>(events), weights,
                       (long long)n_events, sino,
                       n_counted ? c->d_counter : nullptr);
    FIR_CHECK_LAUNCH();
  }
  if (n_counted)
  {
    unsigned long long h = 0;
    FIR_HIP_CHECK(hipMemcpyAsync(&h, c->d_counter, sizeof(h),
                                 hipMemcpyDeviceToHost, c->stream));
    FIR_HIP_CHECK(hipStreamSynchronize(c->stream));
    *n_counted = (int64_t)h;
  }
  return FIR_OK;
}

int fir_rebin_weight(fir_ctx* c, float* sino)
{
  if (!c) return fail(FIR_E_ARG, "context is NULL");
  if (!c->proj_set) return fail(FIR_E_STATE, "sinogram geometry not set");
  if (!sino) return fail(FIR_E_ARG, "NULL sinogram");
  if (c->span == 1) return FIR_OK;
  FIR_HIP_CHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_rebin_weight, dim3(ew_blocks(c->n_bins)), dim3(kBlock),
                     0, c->stream, bin_geom(c), sino);
  FIR_CHECK_LAUNCH();
  return FIR_OK;
}

} // extern "C"
