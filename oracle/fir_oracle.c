/*
 * fir_oracle.c -- CPU restatement of FIR's projector + OSEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the parity checker for the HIP
 * product path in fir_amd/. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product never calls it.
 *
 * It is a plain-C restatement (not a copy) of the reference algorithm,
 * keeping the reference's floating-point types and operation order:
 *   - FP64 geometry and Siddon walk, EPSILON = 1e-5 merge of plane crossings
 *     (reference Siddon.cc:112-412, types.h:10-11, macros.h:8);
 *   - the int-truncated voxel step extents of the reference
 *     (Siddon.h:15-16,113; Siddon.cc:35-38) in "reference" mode, and the
 *     true extents in "geometric" mode;
 *   - FP32 values accumulated as float += double*float
 *     (VolData.cc:357-385);
 *   - the FP32 clamps of VolData.cc:305-343, ProjData.cc:481-539 and
 *     reconAlgos.cc:152-157.
 * Build with -ffp-contract=off so that no multiply-add is fused: the
 * reference is built for plain x86-64 (no FMA), and the walk's decisions
 * depend on every rounding.
 *
 * Every function also exists in an all-FP64 variant (suffix _f64): images
 * and sinograms in double, accumulation in double. That is the "FP64
 * oracle" that errors are reported against.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   orc_scanner_tables   closed form R(cos 2pi i/N, sin 2pi i/N) for one
 *                        crystal per rSector; golden tables from the compiled
 *                        reference.
 *   orc_proj_fill        the reference unit-test known answers
 *                        (ProjHeaderUnitTest.cc:181-232).
 *   orc_crystal_*        closed forms (separation = nViews - |t|) and golden
 *                        LOR tables.
 *   orc_bin_coords       round trip with orc_crystal_* over every bin of
 *                        span-1 and span-3 geometries; the closed forms of
 *                        the tangential coordinate and the view.
 *   orc_rebin_weight     brute-force count of the ring pairs that
 *                        orc_bin_coords fuses into each (segment, axial).
 *   orc_path             the 28 reference unit-test paths
 *                        (SiddonUnitTest.cc:80-423), chord-length closed
 *                        form, brute-force crossing enumeration on tiny
 *                        grids, golden path dumps of the compiled reference.
 *   orc_forward/backward golden sinograms/images of the compiled reference;
 *                        adjoint identity <Ax,y> = <x,A^T y> in FP64.
 *   orc_osem             golden images of the compiled reference.
 *   orc_osem_reso        golden images of the compiled reference's
 *                        reconAlgos::OSEM_ResoReco (called directly: no
 *                        executable reaches it).
 *   orc_convolve         delta-image closed form; golden output.
 *   orc_cut_circle, orc_hu_to_mu, orc_proj_exp, orc_*_mul/div
 *                        closed forms on hand-built inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EPS 0.00001
#define ORC_MAX_SEG 1024

/* ------------------------------------------------------------------------ */
/* Geometry descriptors (plain structs shared with ctypes)                   */
/* ------------------------------------------------------------------------ */

typedef struct
{
  int nx, ny, nz;
  double dx, dy, dz; /* voxel extent, mm */
  double ox, oy, oz; /* x,y: centre of first voxel; z: offset of centre */
} orc_vol;

typedef struct
{
  int n_rings, n_cpr, span, n_seg, n_tang;
  int seg_off, tang_off, n_views, half_span, max_ring_diff;
  long long n_bins;
  int n_axial[ORC_MAX_SEG];
  long long seg_base[ORC_MAX_SEG]; /* first bin of each segment */
} orc_proj;

/* Derived walk constants (reference Siddon ctor, Siddon.cc:20-52). */
typedef struct
{
  double low[3], high[3];
  double step[3]; /* step extent used by the walk */
  int nm1[3];
  int row, slice;
  int max_path;
} orc_grid;

/* ------------------------------------------------------------------------ */
/* Projection geometry: ProjGeometry::fill (ProjHeader.cc:108-186)           */
/* ------------------------------------------------------------------------ */

/* Returns 0 on success, nonzero if the header is invalid (the reference
 * ProjHeader::check conditions, ProjHeader.cc:6-106). */
int orc_proj_fill(orc_proj* p, int n_rings, int n_cpr, int span, int n_seg,
                  int n_tang)
{
  if (n_rings <= 0 || n_cpr <= 0 || n_cpr % 4 != 0) return 1;
  if (span <= 0 || span % 2 != 1 || span > 2 * n_rings - 1) return 2;
  if (n_seg < 0 || n_seg % 2 != 1) return 3;
  {
    int max_seg = (2 * n_rings - 1) / span;
    max_seg -= 1 - max_seg % 2;
    if (n_seg > max_seg) return 3;
  }
  if (n_seg > ORC_MAX_SEG) return 3;
  if (n_tang < 0 || n_tang > n_cpr - 1) return 4;
  if (n_tang == 0) n_tang = n_cpr - 1;

  memset(p, 0, sizeof(*p));
  p->n_rings = n_rings;
  p->n_cpr = n_cpr;
  p->span = span;
  p->n_seg = n_seg;
  p->n_tang = n_tang;
  p->seg_off = (n_seg - 1) / 2;
  p->tang_off = n_tang / 2;
  p->n_views = n_cpr / 2;
  p->half_span = (span - 1) / 2;

  for (int s = -p->seg_off; s <= p->seg_off; ++s)
  {
    int a = s < 0 ? -s : s;
    int len;
    if (span == 1)
      len = n_rings - a;
    else
    {
      len = 2 * n_rings - 1;
      if (a >= 1) len -= span + 1;
      if (a >= 2) len -= 2 * span * (a - 1);
    }
    p->n_axial[s + p->seg_off] = len;
  }

  long long acc = 0;
  for (int k = 0; k < n_seg; ++k)
  {
    p->seg_base[k] = acc;
    acc += (long long)p->n_axial[k] * p->n_views * n_tang;
  }
  p->n_bins = acc;

  if (span == 1)
    p->max_ring_diff = p->seg_off;
  else
    p->max_ring_diff = p->half_span + (p->seg_off > 0 ? p->seg_off * span : 0);
  return 0;
}

/* (seg, axial) -> slice indices of the two crystals
 * (reference ProjData::getCrystalAxialCoord, ProjData.cc:198-299). */
void orc_crystal_axial(const orc_proj* p, int seg, int axial, int* s1, int* s2)
{
  if (p->span == 1)
  {
    if (seg == 0)
    {
      *s1 = *s2 = 2 * axial;
    }
    else if (seg > 0)
    {
      *s1 = 2 * (axial + seg);
      *s2 = 2 * axial;
    }
    else
    {
      *s1 = 2 * axial;
      *s2 = 2 * (axial - seg);
    }
    return;
  }
  if (seg == 0)
  {
    *s1 = *s2 = axial;
    return;
  }
  const int aseg = seg < 0 ? -seg : seg;
  const int mid = aseg * p->span;
  const int hs = p->half_span;
  if (axial < hs)
  {
    const int rd = mid - hs + axial;
    if (seg > 0) { *s1 = 2 * rd; *s2 = 0; }
    else { *s1 = 0; *s2 = 2 * rd; }
    return;
  }
  const int inv = p->n_axial[seg + p->seg_off] - axial - 1;
  if (inv < hs)
  {
    const int rd = mid - hs + inv;
    const int smax = 2 * p->n_rings - 2;
    if (seg > 0) { *s1 = smax; *s2 = smax - 2 * rd; }
    else { *s1 = smax - 2 * rd; *s2 = smax; }
    return;
  }
  const int adj = axial - hs;
  if (seg > 0) { *s1 = 2 * mid + adj; *s2 = adj; }
  else { *s1 = adj; *s2 = 2 * mid + adj; }
}

/* (view, tang) -> angular crystal indices
 * (reference ProjData::getCrystalAngCoord, ProjData.cc:301-340).
 * Integer division truncates toward zero, as in C++. */
void orc_crystal_ang(const orc_proj* p, int view, int tang, int* c1, int* c2)
{
  const int shift = tang >= 0 ? tang / 2 : (tang - 1) / 2;
  const int odd = (tang < 0 ? -tang : tang) % 2;
  int a = view + shift + odd;
  int b = view + p->n_views - shift;
  const int n = p->n_cpr;
  if (a < 0) a += n;
  else if (a >= n) a -= n;
  if (b < 0) b += n;
  else if (b >= n) b -= n;
  *c1 = a;
  *c2 = b;
}

/* Crystal pair -> sinogram bin coordinates (reference
 * ProjData::getBinCoordinates, ProjData.cc:342-479). r = ring (axial
 * crystal index), c = angular crystal index. Returns 0 when the pair lies
 * outside the sinogram (ring difference above the largest one stored, or
 * tangential coordinate outside the stored range), else 1 with
 * (seg, view, axial, tang). C integer division truncates toward zero, as
 * C++'s does. */
int orc_bin_coords(const orc_proj* p, int r1, int c1, int r2, int c2,
                   int* seg, int* view, int* axial, int* tang)
{
  const int n = p->n_cpr;
  int aseg = r1 > r2 ? r1 - r2 : r2 - r1;
  if (aseg > p->max_ring_diff) return 0;
  /* tangential: the crystal sum decides which half-turn the LOR is in */
  const int sum = c1 + c2;
  const int s1 = (sum >= n / 2 && sum < 3 * n / 2) ? -1 : 1;
  const int dc = c2 > c1 ? c2 - c1 : c1 - c2;
  const int t = s1 * (dc - n / 2);
  if (t < -p->tang_off || t >= -p->tang_off + p->n_tang) return 0;
  /* axial: span 1 = lower ring; span > 1: ring sum minus the first ring
   * sum of the (fused) segment */
  int ax;
  if (p->span == 1)
    ax = (r1 + r2 - aseg) / 2;
  else
  {
    const int central = aseg <= p->half_span;
    aseg = central ? 0 : 1 + (aseg - p->half_span - 1) / p->span;
    const int first = central ? 0 : 1 + p->half_span + (aseg - 1) * p->span;
    ax = r1 + r2 - first;
  }
  const int v = ((sum + n / 2) % n) / 2;
  int sg = aseg;
  if (sg != 0)
  {
    /* the sign follows from the pair's crystals moved to the parallel LOR
     * through the centre (tangential 0 or 1) */
    const int s2 = c1 < c2 ? 1 : -1;
    const int at = t < 0 ? -t : t;
    const int u = at % 2 == 0 ? -t / 2 : -(t - 1) / 2;
    int k1 = c1 - s1 * s2 * u;
    int k2 = c2 + s1 * s2 * u;
    if (k1 >= n) k1 -= n;
    else if (k1 < 0) k1 += n;
    if (k2 >= n) k2 -= n;
    else if (k2 < 0) k2 += n;
    const int ring_sign = r1 < r2 ? -1 : 1;
    const int flip = k1 < k2 ? 1 : -1;
    sg = sg * ring_sign * flip;
  }
  *seg = sg;
  *view = v;
  *axial = ax;
  *tang = t;
  return 1;
}

/* Flat bin index of (seg, view, axial, tang) in file order, or -1 when the
 * coordinates are outside the sinogram. */
long long orc_bin_index(const orc_proj* p, int seg, int view, int axial,
                        int tang)
{
  const int k = seg + p->seg_off;
  if (k < 0 || k >= p->n_seg || view < 0 || view >= p->n_views) return -1;
  if (axial < 0 || axial >= p->n_axial[k]) return -1;
  const int t = tang + p->tang_off;
  if (t < 0 || t >= p->n_tang) return -1;
  return p->seg_base[k] +
    ((long long)view * p->n_axial[k] + axial) * p->n_tang + t;
}

/* Histogram of list-mode events (4 ints each: r1, c1, r2, c2) into the
 * sinogram, one count (or weight w[i]) per event whose pair maps to a bin
 * (getBinCoordinates + incrementBin, ProjData.cc:342-479, ProjData.inl:47-56).
 * Events with crystal indices out of range are skipped. Returns the number
 * of events counted. */
long long orc_histogram_events(const orc_proj* p, const int* ev, long long n,
                               const float* w, float* sino)
{
  long long counted = 0;
  for (long long i = 0; i < n; ++i)
  {
    const int* e = ev + 4 * i;
    if (e[0] < 0 || e[0] >= p->n_rings || e[2] < 0 || e[2] >= p->n_rings ||
        e[1] < 0 || e[1] >= p->n_cpr || e[3] < 0 || e[3] >= p->n_cpr)
      continue;
    int sg, v, ax, t;
    if (!orc_bin_coords(p, e[0], e[1], e[2], e[3], &sg, &v, &ax, &t))
      continue;
    const long long b = orc_bin_index(p, sg, v, ax, t);
    if (b < 0) continue;
    sino[b] += w ? w[i] : 1.0f;
    ++counted;
  }
  return counted;
}

/* Span > 1: every bin of a (segment, axial) position that fuses k > 1 ring
 * pairs is multiplied by (float)(1.0 / k) (reference ProjData::rebinWeight,
 * ProjData.cc:541-615, walking ring sums and, within each, ring differences
 * of its parity segment by segment). Bins are addressed with the correct
 * per-segment axial count (the reference's ProjData::getInd,
 * ProjData.inl:67-82, offsets the segment twice). No-op for span 1. */
void orc_rebin_weight(const orc_proj* p, float* sino)
{
  if (p->span == 1) return;
  const int nr = p->n_rings;
  for (int rs = 0; rs <= 2 * (nr - 1); ++rs)
  {
    int lmax = rs <= nr - 1 ? rs : 2 * (nr - 1) - rs;
    if (lmax > p->max_ring_diff) lmax = p->max_ring_diff;
    if ((rs + lmax) % 2) --lmax;
    int sg = -p->seg_off - 1;
    int next = -p->max_ring_diff - 1;
    while (next < -lmax)
    {
      ++sg;
      next += p->span;
    }
    int k = 1;
    for (int d = -lmax; d <= lmax; d += 2)
    {
      if (d == lmax || d + 1 >= next)
      {
        if (k > 1)
        {
          const int asg = sg < 0 ? -sg : sg;
          const int ax = sg == 0 ? rs
                                 : rs - (p->span + 1) / 2 - (asg - 1) * p->span;
          const float wt = (float)(1.0 / k);
          for (int v = 0; v < p->n_views; ++v)
            for (int t = -p->tang_off; t < -p->tang_off + p->n_tang; ++t)
            {
              const long long b = orc_bin_index(p, sg, v, ax, t);
              if (b >= 0) sino[b] *= wt;
            }
        }
        k = 1;
        ++sg;
        next += p->span;
      }
      else
        ++k;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Scanner tables (ScannerGeometry::fill + ScannerData position vectors,     */
/* ScannerHeader.cc:305-342, ScannerData.cc:234-377)                        */
/* ------------------------------------------------------------------------ */

/* hdr layout (doubles):
 *  [0..2] crystal dims XYZ, [3..4] crystal repeat YZ, [5..6] inter-crystal YZ,
 *  [7..9] module dims XYZ, [10..11] module repeat YZ, [12..13] inter-module YZ,
 *  [14..16] rSector dims XYZ, [17] rSector repeat, [18] inner radius.
 * Zero module / rSector dimensions are replaced by their tight minimum, as
 * ScannerHeader::check does. xy gets 2*n_cpr doubles, z gets 2*n_rings-1. */
int orc_scanner_tables(const double* hdr, double* xy, double* z,
                       int* n_rings_out, int* n_cpr_out)
{
  double cd[3] = {hdr[0], hdr[1], hdr[2]};
  int crep[2] = {(int)hdr[3], (int)hdr[4]};
  double cgap[2] = {hdr[5], hdr[6]};
  double md[3] = {hdr[7], hdr[8], hdr[9]};
  int mrep[2] = {(int)hdr[10], (int)hdr[11]};
  double mgap[2] = {hdr[12], hdr[13]};
  double rd[3] = {hdr[14], hdr[15], hdr[16]};
  int rrep = (int)hdr[17];
  double rin = hdr[18];

  if (md[0] == 0.0) md[0] = cd[0];
  if (md[1] == 0.0) md[1] = cd[1] * crep[0] + cgap[0] * (crep[0] - 1);
  if (md[2] == 0.0) md[2] = cd[2] * crep[1] + cgap[1] * (crep[1] - 1);
  if (rd[0] == 0) rd[0] = md[0];
  if (rd[1] == 0) rd[1] = md[1] * mrep[0] + mgap[0] * (mrep[0] - 1);
  if (rd[2] == 0) rd[2] = md[2] * mrep[1] + mgap[1] * (mrep[1] - 1);

  const double cvec_y = cd[1] + cgap[0];
  const double cvec_z = cd[2] + cgap[1];
  const double mvec_y = md[1] + mgap[0];
  const double mvec_z = md[2] + mgap[1];
  const double tx = rin + rd[0] / 2;
  const int n_cpr = crep[0] * mrep[0] * rrep;
  const int n_rings = crep[1] * mrep[1];
  const int coff = mrep[0] * crep[0] / 2;

  if (xy)
  {
    const int per_sector = mrep[0] * crep[0];
    double* ys = (double*)malloc(sizeof(double) * per_sector);
    const double y0 = -((mrep[0] - 1) * mvec_y + (crep[0] - 1) * cvec_y) / 2.0;
    for (int m = 0; m < mrep[0]; ++m)
      for (int c = 0; c < crep[0]; ++c)
        ys[m * crep[0] + c] = m * mvec_y + c * cvec_y + y0;
    const double dang = 2 * 3.141592653589793 / rrep;
    const int first = n_cpr - coff;
    for (int i = 0; i < first; ++i)
    {
      const int k = i + coff;
      const double yy = ys[k % per_sector];
      const double ang = (k / per_sector) * dang;
      const double ca = cos(ang), sa = sin(ang);
      xy[2 * i] = tx * ca - yy * sa;
      xy[2 * i + 1] = tx * sa + yy * ca;
    }
    for (int i = first; i < n_cpr; ++i)
    {
      xy[2 * i] = tx;
      xy[2 * i + 1] = ys[i - first];
    }
    free(ys);
  }
  if (z)
  {
    const int per = mrep[1] * crep[1];
    double* rz = (double*)malloc(sizeof(double) * per);
    const double z0 = -((mrep[1] - 1) * mvec_z + (crep[1] - 1) * cvec_z) / 2.0;
    for (int m = 0; m < mrep[1]; ++m)
      for (int c = 0; c < crep[1]; ++c)
        rz[m * crep[1] + c] = m * mvec_z + c * cvec_z + z0;
    const int ns = 2 * n_rings - 1;
    for (int s = 0; s < ns; ++s)
      z[s] = (s % 2 == 0) ? rz[s / 2] : (rz[s / 2] + rz[(s + 1) / 2]) / 2.0;
    free(rz);
  }
  if (n_rings_out) *n_rings_out = n_rings;
  if (n_cpr_out) *n_cpr_out = n_cpr;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Siddon walk (Siddon.cc:112-412)                                          */
/* ------------------------------------------------------------------------ */

static void grid_init(orc_grid* g, const orc_vol* v, int geometric)
{
  const double ext[3] = {v->dx, v->dy, v->dz};
  const int n[3] = {v->nx, v->ny, v->nz};
  g->low[0] = v->ox - v->dx / 2;
  g->low[1] = v->oy - v->dy / 2;
  g->low[2] = v->oz - (v->nz * v->dz) / 2;
  for (int d = 0; d < 3; ++d)
  {
    g->high[d] = g->low[d] + n[d] * ext[d];
    g->step[d] = geometric ? ext[d] : (double)(int)ext[d];
    g->nm1[d] = n[d] - 1;
  }
  g->row = v->nx;
  g->slice = v->nx * v->ny;
  g->max_path = v->nx + v->ny + v->nz;
}

#define ORC_MIN(a, b) ((a) < (b) ? (a) : (b))
#define ORC_MAX(a, b) ((a) > (b) ? (a) : (b))
#define ORC_ABS(a) ((a) < 0 ? (-(a)) : (a))

/* Returns the number of path elements, or -1 if the line misses the
 * volume (the reference returns false). At most cap elements are written;
 * the walk itself is bounded by max_path + 8 steps. */
static int grid_path(const orc_grid* g, const double* c1, const double* c2,
                     int* idx, double* len, int cap)
{
  double diff[3], amin[3], amax[3], nxt[3];
  int dir[3];
  for (int d = 0; d < 3; ++d)
  {
    double df = c2[d] - c1[d];
    if (ORC_ABS(df) > ORC_EPS)
    {
      if (df > 0)
      {
        dir[d] = 1;
        amin[d] = (g->low[d] - c1[d]) / df;
        amax[d] = (g->high[d] - c1[d]) / df;
      }
      else
      {
        dir[d] = -1;
        amin[d] = (g->high[d] - c1[d]) / df;
        amax[d] = (g->low[d] - c1[d]) / df;
      }
      nxt[d] = 0.0;
    }
    else
    {
      if (c1[d] < g->low[d] || c1[d] > g->high[d]) return -1;
      df = ORC_EPS;
      dir[d] = -1;
      amin[d] = 0.0;
      amax[d] = 1.0;
      nxt[d] = 1.0;
    }
    diff[d] = df;
  }
  const double a_lo = ORC_MAX(amin[0], ORC_MAX(amin[1], ORC_MAX(amin[2], 0.0)));
  const double a_hi = ORC_MIN(amax[0], ORC_MIN(amax[1], ORC_MIN(amax[2], 1.0)));
  if (a_lo >= a_hi) return -1;

  double da[3], a[3];
  int pos[3];
  for (int d = 0; d < 3; ++d) da[d] = g->step[d] / ORC_ABS(diff[d]);
  const double d12 =
    sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
  for (int d = 0; d < 3; ++d)
  {
    int i = (int)((c1[d] + diff[d] * a_lo - g->low[d]) / g->step[d]);
    i = ORC_MAX(0, ORC_MIN(i, g->nm1[d]));
    pos[d] = i;
  }
  for (int d = 0; d < 3; ++d)
  {
    double al = nxt[d];
    if (al < 1.0) al = (g->low[d] + g->step[d] * pos[d] - c1[d]) / diff[d];
    if (dir[d] > 0) al += da[d];
    a[d] = al;
  }

  int n = 0;
  double prev = a_lo;
  const int limit = g->max_path + 8;
  for (int it = 0; prev < a_hi && it < limit; ++it)
  {
    const double next = ORC_MIN(a_hi, ORC_MIN(a[0], ORC_MIN(a[1], a[2])));
    if (pos[0] >= 0 && pos[0] <= g->nm1[0] && pos[1] >= 0 &&
        pos[1] <= g->nm1[1] && pos[2] >= 0 && pos[2] <= g->nm1[2])
    {
      if (n < cap)
      {
        idx[n] = pos[0] + pos[1] * g->row + pos[2] * g->slice;
        len[n] = (next - prev) * d12;
      }
      ++n;
    }
    for (int d = 0; d < 3; ++d)
    {
      if (ORC_ABS(a[d] - next) < ORC_EPS)
      {
        a[d] += da[d];
        pos[d] += dir[d];
      }
    }
    prev = next;
  }
  return n;
}

/* Path of one line given its two endpoints. idx/len need cap entries. */
int orc_path(const orc_vol* v, int geometric, const double* p1,
             const double* p2, int* idx, double* len, int cap)
{
  orc_grid g;
  grid_init(&g, v, geometric);
  return grid_path(&g, p1, p2, idx, len, cap);
}

/* Endpoints of the LOR in bin (seg, view, axial, tang). */
static void lor_endpoints(const orc_proj* p, const double* xy, const double* z,
                          int seg, int view, int axial, int tang, double* e1,
                          double* e2)
{
  int s1, s2, a1, a2;
  orc_crystal_axial(p, seg, axial, &s1, &s2);
  orc_crystal_ang(p, view, tang, &a1, &a2);
  e1[0] = xy[2 * a1];
  e1[1] = xy[2 * a1 + 1];
  e1[2] = z[s1];
  e2[0] = xy[2 * a2];
  e2[1] = xy[2 * a2 + 1];
  e2[2] = z[s2];
}

/* Decode the linear bin index b (segment-concatenated order) into
 * (seg, view, axial, tang). */
static void bin_decode(const orc_proj* p, long long b, int* seg, int* view,
                       int* axial, int* tang)
{
  int k = 0;
  while (k + 1 < p->n_seg && p->seg_base[k + 1] <= b) ++k;
  long long r = b - p->seg_base[k];
  const int na = p->n_axial[k];
  const long long per_view = (long long)na * p->n_tang;
  *seg = k - p->seg_off;
  *view = (int)(r / per_view);
  r -= (long long)(*view) * per_view;
  *axial = (int)(r / p->n_tang);
  *tang = (int)(r - (long long)(*axial) * p->n_tang) - p->tang_off;
}

/* All paths in bin order: counts[b] (= -1 for a miss) and the concatenated
 * elements of bins [b0, b0+nb). Returns the total number of elements, which
 * may exceed cap (then only the first cap are written). */
long long orc_path_dump(const orc_vol* v, int geometric, const orc_proj* p,
                        const double* xy, const double* z, long long b0,
                        long long nb, int* counts, int* idx, double* len,
                        long long cap)
{
  orc_grid g;
  grid_init(&g, v, geometric);
  int* ti = (int*)malloc(sizeof(int) * (g.max_path + 16));
  double* tl = (double*)malloc(sizeof(double) * (g.max_path + 16));
  long long tot = 0;
  for (long long b = b0; b < b0 + nb; ++b)
  {
    int seg, view, axial, tang;
    double e1[3], e2[3];
    bin_decode(p, b, &seg, &view, &axial, &tang);
    lor_endpoints(p, xy, z, seg, view, axial, tang, e1, e2);
    const int n = grid_path(&g, e1, e2, ti, tl, g.max_path + 16);
    counts[b - b0] = n;
    for (int e = 0; e < n; ++e)
    {
      if (tot + e < cap)
      {
        idx[tot + e] = ti[e];
        len[tot + e] = tl[e];
      }
    }
    if (n > 0) tot += n;
  }
  free(ti);
  free(tl);
  return tot;
}

/* ------------------------------------------------------------------------ */
/* Projectors, OSEM and image operations, instantiated for float (reference  */
/* precision) and double (all-FP64 oracle).                                 */
/* ------------------------------------------------------------------------ */

static int set_threads(int nthreads)
{
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  return omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

/* Gaussian kernel construction of operations::convolve
 * (operations.cc:31-75): X and Y use a float sigma, Z a double sigma. The
 * taps are computed in double and stored as float. Writes 2*h+1 taps per
 * axis into k{x,y,z} (capacity 64 each) and the half widths into h[3]. */
int orc_gauss_taps(const orc_vol* v, const float* fwhm, float* kx, float* ky,
                   float* kz, int* h)
{
  const float sx = fwhm[0] / 2.3548f;
  const float sy = fwhm[1] / 2.3548f;
  const double sz = fwhm[2] / 2.3548;
  h[0] = ((int)(6.0 * sx / v->dx) + 1) / 2;
  h[1] = ((int)(6.0 * sy / v->dy) + 1) / 2;
  h[2] = ((int)(6.0 * sz / v->dz) + 1) / 2;
  if (h[0] > 31 || h[1] > 31 || h[2] > 31) return 1;
  for (int k = -h[0]; k <= h[0]; ++k)
    kx[h[0] + k] = (float)exp(-(k * k * v->dx * v->dx) / (2 * sx * sx));
  for (int k = -h[1]; k <= h[1]; ++k)
    ky[h[1] + k] = (float)exp(-(k * k * v->dy * v->dy) / (2 * sy * sy));
  for (int k = -h[2]; k <= h[2]; ++k)
    kz[h[2] + k] = (float)exp(-(k * k * v->dz * v->dz) / (2 * sz * sz));
  return 0;
}

/* Mu-map conversion (operations.cc:243-284), float semantics. */
void orc_hu_to_mu(float* img, long long n)
{
  const float water = 0.0096f;
  const float thousand = 0.015f;
  const float s1 = water / (float)1000.0;
  const float s2 = (thousand - water) / (float)1000.0;
  for (long long i = 0; i < n; ++i)
  {
    if (img[i] <= -1000.0)
      img[i] = 0.0f;
    else if (img[i] <= 0.0)
    {
      img[i] *= s1;
      img[i] += water;
    }
    else
    {
      img[i] *= s2;
      img[i] += water;
    }
  }
}

#define ORC_REAL float
#define ORC_SFX(name) name
#include "fir_oracle_body.h"
#undef ORC_REAL
#undef ORC_SFX

#define ORC_REAL double
#define ORC_SFX(name) name##_f64
#include "fir_oracle_body.h"
#undef ORC_REAL
#undef ORC_SFX

int orc_sizeof_proj(void) { return (int)sizeof(orc_proj); }
int orc_sizeof_vol(void) { return (int)sizeof(orc_vol); }
