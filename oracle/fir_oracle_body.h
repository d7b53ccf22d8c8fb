/*
 * fir_oracle_body.h -- value-type-generic part of the CPU oracle.
 * Included twice by fir_oracle.c: ORC_REAL = float (reference precision)
 * and ORC_REAL = double (all-FP64 oracle, names suffixed _f64).
 * TEST INFRASTRUCTURE ONLY (see fir_oracle.c).
 */

/* Forward projection of every bin (projections.cc:15-92): per segment, an
 * OpenMP loop over views; each bin gets sum(len * img) accumulated in path
 * order into an ORC_REAL (float += double*float for the float variant).
 * Bins whose LOR misses the volume get 0. */
void ORC_SFX(orc_forward)(const orc_vol* v, int geometric, const orc_proj* p,
                          const double* xy, const double* z,
                          const ORC_REAL* img, ORC_REAL* sino, int nthreads)
{
  orc_grid g;
  grid_init(&g, v, geometric);
  set_threads(nthreads);
  for (int seg = -p->seg_off; seg <= p->seg_off; ++seg)
  {
    const int na = p->n_axial[seg + p->seg_off];
    ORC_REAL* out = sino + p->seg_base[seg + p->seg_off];
#pragma omp parallel
    {
      int* ti = (int*)malloc(sizeof(int) * (g.max_path + 16));
      double* tl = (double*)malloc(sizeof(double) * (g.max_path + 16));
#pragma omp for schedule(static)
      for (int view = 0; view < p->n_views; ++view)
      {
        long long b = (long long)view * na * p->n_tang;
        for (int axial = 0; axial < na; ++axial)
          for (int t = -p->tang_off; t < -p->tang_off + p->n_tang; ++t)
          {
            double e1[3], e2[3];
            lor_endpoints(p, xy, z, seg, view, axial, t, e1, e2);
            const int n = grid_path(&g, e1, e2, ti, tl, g.max_path + 16);
            ORC_REAL line = 0;
            for (int e = 0; e < n; ++e) line += tl[e] * img[ti[e]];
            out[b++] = line;
          }
      }
      free(ti);
      free(tl);
    }
  }
}

/* Forward projection of the listed bins only (all-bins order): out[i] is
 * what orc_forward writes for bin bins[i], with the same accumulation. */
void ORC_SFX(orc_forward_bins)(const orc_vol* v, int geometric,
                               const orc_proj* p, const double* xy,
                               const double* z, const ORC_REAL* img,
                               const long long* bins, long long n,
                               ORC_REAL* out, int nthreads)
{
  orc_grid g;
  grid_init(&g, v, geometric);
  set_threads(nthreads);
#pragma omp parallel
  {
    int* ti = (int*)malloc(sizeof(int) * (g.max_path + 16));
    double* tl = (double*)malloc(sizeof(double) * (g.max_path + 16));
#pragma omp for schedule(dynamic, 256)
    for (long long i = 0; i < n; ++i)
    {
      int seg, view, axial, tang;
      double e1[3], e2[3];
      bin_decode(p, bins[i], &seg, &view, &axial, &tang);
      lor_endpoints(p, xy, z, seg, view, axial, tang, e1, e2);
      const int m = grid_path(&g, e1, e2, ti, tl, g.max_path + 16);
      ORC_REAL line = 0;
      for (int e = 0; e < m; ++e) line += tl[e] * img[ti[e]];
      out[i] = line;
    }
    free(ti);
    free(tl);
  }
}

/* Backprojection of the listed bins only, added into img (not cleared):
 * each bin's path scattered with vals[i], as orc_backward scatters it. */
void ORC_SFX(orc_backward_bins)(const orc_vol* v, int geometric,
                                const orc_proj* p, const double* xy,
                                const double* z, const long long* bins,
                                const ORC_REAL* vals, long long n,
                                ORC_REAL* img, int nthreads)
{
  orc_grid g;
  grid_init(&g, v, geometric);
  set_threads(nthreads);
#pragma omp parallel
  {
    int* ti = (int*)malloc(sizeof(int) * (g.max_path + 16));
    double* tl = (double*)malloc(sizeof(double) * (g.max_path + 16));
#pragma omp for schedule(dynamic, 256)
    for (long long i = 0; i < n; ++i)
    {
      int seg, view, axial, tang;
      double e1[3], e2[3];
      bin_decode(p, bins[i], &seg, &view, &axial, &tang);
      lor_endpoints(p, xy, z, seg, view, axial, tang, e1, e2);
      const int m = grid_path(&g, e1, e2, ti, tl, g.max_path + 16);
      const ORC_REAL val = vals[i];
      for (int e = 0; e < m; ++e)
      {
#pragma omp atomic
        img[ti[e]] += tl[e] * val;
      }
    }
    free(ti);
    free(tl);
  }
}

/* Visit the LORs of (subset, seg) in the reference LOR-cache order
 * (LORCache.cc:49-51,118-133): index -> (subview, axial, tang) with
 * view = subview * nSubsets + subset. */
#ifndef ORC_SUBSET_DECODE
#define ORC_SUBSET_DECODE
static inline void subset_decode(const orc_proj* p, int na, int n_subsets,
                                 int subset, long long index, int* view,
                                 int* axial, int* tang, long long* bin)
{
  const long long per_view = (long long)na * p->n_tang;
  const long long sv = index / per_view;
  const long long rem = index - sv * per_view;
  *view = (int)(sv * n_subsets + subset);
  *axial = (int)(rem / p->n_tang);
  *tang = (int)(rem - (long long)(*axial) * p->n_tang) - p->tang_off;
  *bin = (long long)(*view) * per_view + rem;
}
#endif

/* Backprojection (projections.cc:94-184): all frames zeroed, then frame s
 * receives the scatter of the LORs of subset s. img has n_subsets frames. */
void ORC_SFX(orc_backward)(const orc_vol* v, int geometric, const orc_proj* p,
                           const double* xy, const double* z,
                           const ORC_REAL* sino, ORC_REAL* img, int n_subsets,
                           int nthreads)
{
  orc_grid g;
  grid_init(&g, v, geometric);
  set_threads(nthreads);
  const long long nvox = (long long)v->nx * v->ny * v->nz;
  memset(img, 0, sizeof(ORC_REAL) * nvox * n_subsets);
  const int vps = p->n_views / n_subsets;
  for (int subset = 0; subset < n_subsets; ++subset)
  {
    ORC_REAL* frame = img + nvox * subset;
    for (int seg = -p->seg_off; seg <= p->seg_off; ++seg)
    {
      const int na = p->n_axial[seg + p->seg_off];
      const ORC_REAL* in = sino + p->seg_base[seg + p->seg_off];
      const long long nlor = (long long)vps * na * p->n_tang;
#pragma omp parallel
      {
        int* ti = (int*)malloc(sizeof(int) * (g.max_path + 16));
        double* tl = (double*)malloc(sizeof(double) * (g.max_path + 16));
#pragma omp for schedule(static)
        for (long long index = 0; index < nlor; ++index)
        {
          int view, axial, tang;
          long long bin;
          double e1[3], e2[3];
          subset_decode(p, na, n_subsets, subset, index, &view, &axial, &tang,
                        &bin);
          lor_endpoints(p, xy, z, seg, view, axial, tang, e1, e2);
          const int n = grid_path(&g, e1, e2, ti, tl, g.max_path + 16);
          const ORC_REAL val = in[bin];
          for (int e = 0; e < n; ++e)
          {
#pragma omp atomic
            frame[ti[e]] += tl[e] * val;
          }
        }
        free(ti);
        free(tl);
      }
    }
  }
}

/* Sensitivity = backprojection of a sinogram of ones, one frame per subset
 * (projections.cc:186-197). */
void ORC_SFX(orc_sensitivity)(const orc_vol* v, int geometric,
                              const orc_proj* p, const double* xy,
                              const double* z, ORC_REAL* sens, int n_subsets,
                              int nthreads)
{
  ORC_REAL* ones = (ORC_REAL*)malloc(sizeof(ORC_REAL) * p->n_bins);
  for (long long b = 0; b < p->n_bins; ++b) ones[b] = 1;
  ORC_SFX(orc_backward)(v, geometric, p, xy, z, ones, sens, n_subsets,
                        nthreads);
  free(ones);
}

/* Clamped voxel/bin arithmetic (VolData.cc:305-343, ProjData.cc:481-539):
 * a = (a > eps && b > eps) ? a op b : 0. */
void ORC_SFX(orc_mul_clamped)(ORC_REAL* a, const ORC_REAL* b, long long n)
{
  for (long long i = 0; i < n; ++i)
    a[i] = (a[i] > ORC_EPS && b[i] > ORC_EPS) ? a[i] * b[i] : 0;
}

void ORC_SFX(orc_div_clamped)(ORC_REAL* a, const ORC_REAL* b, long long n)
{
  for (long long i = 0; i < n; ++i)
    a[i] = (a[i] > ORC_EPS && b[i] > ORC_EPS) ? a[i] / b[i] : 0;
}

/* ACF exponential (ProjData.cc:514-539): exp(x) where x > eps, else 1. */
void ORC_SFX(orc_exp_clamped)(ORC_REAL* a, long long n)
{
  for (long long i = 0; i < n; ++i)
    a[i] = a[i] > ORC_EPS ? (sizeof(ORC_REAL) == 4 ? expf((float)a[i])
                                                   : exp((double)a[i]))
                          : 1;
}

/* Circular cut (operations.cc:194-220): zero voxels whose centre, relative
 * to the image centre, lies farther than cut from the axis. */
void ORC_SFX(orc_cut_circle)(const orc_vol* v, ORC_REAL* img, float cut)
{
  if (!(cut > 0.0)) return;
  const double sw = v->nx * v->dx, sh = v->ny * v->dy;
  for (int k = 0; k < v->nz; ++k)
    for (int j = 0; j < v->ny; ++j)
      for (int i = 0; i < v->nx; ++i)
      {
        const double px = i * v->dx + v->dx / 2.0 - sw / 2.0;
        const double py = j * v->dy + v->dy / 2.0 - sh / 2.0;
        if (sqrt(px * px + py * py) > cut)
          img[i + (long long)j * v->nx + (long long)k * v->nx * v->ny] = 0;
      }
}

/* Separable Gaussian with border renormalisation and optional edge restore
 * (operations.cc:11-192). One frame, in place. */
void ORC_SFX(orc_convolve)(const orc_vol* v, ORC_REAL* img, const float* fwhm,
                           float cut)
{
  if (!(fwhm[0] > 0.0 && fwhm[1] > 0.0 && fwhm[2] > 0.0)) return;
  float kx[64], ky[64], kz[64];
  int h[3];
  if (orc_gauss_taps(v, fwhm, kx, ky, kz, h)) return;
  const int nx = v->nx, ny = v->ny, nz = v->nz;
  const long long n = (long long)nx * ny * nz;
  ORC_REAL* keep = (ORC_REAL*)malloc(sizeof(ORC_REAL) * n);
  ORC_REAL* t1 = (ORC_REAL*)malloc(sizeof(ORC_REAL) * n);
  ORC_REAL* t2 = (ORC_REAL*)malloc(sizeof(ORC_REAL) * n);
  memcpy(keep, img, sizeof(ORC_REAL) * n);
#define ORC_AT(i, j, k) ((i) + (long long)(j) * nx + (long long)(k) * nx * ny)
#pragma omp parallel for
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
      {
        ORC_REAL s = 0, w = 0;
        for (int q = -h[0]; q <= h[0]; ++q)
          if (i + q >= 0 && i + q < nx)
          {
            const ORC_REAL kv = kx[h[0] + q];
            s += kv * img[ORC_AT(i + q, j, k)];
            w += kv;
          }
        if (w > 0.0) t1[ORC_AT(i, j, k)] = s / w;
      }
#pragma omp parallel for
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
      {
        ORC_REAL s = 0, w = 0;
        for (int q = -h[1]; q <= h[1]; ++q)
          if (j + q >= 0 && j + q < ny)
          {
            const ORC_REAL kv = ky[h[1] + q];
            s += kv * t1[ORC_AT(i, j + q, k)];
            w += kv;
          }
        if (w > 0.0) t2[ORC_AT(i, j, k)] = s / w;
      }
#pragma omp parallel for
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
      {
        ORC_REAL s = 0, w = 0;
        for (int q = -h[2]; q <= h[2]; ++q)
          if (k + q >= 0 && k + q < nz)
          {
            const ORC_REAL kv = kz[h[2] + q];
            s += kv * t2[ORC_AT(i, j, k + q)];
            w += kv;
          }
        if (w > 0.0) img[ORC_AT(i, j, k)] = s / w;
      }
  if (cut > 0.0)
  {
    const float fw = fwhm[0] > fwhm[1] ? fwhm[0] : fwhm[1];
    const double sw = nx * v->dx, sh = ny * v->dy;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
        {
          const double px = i * v->dx + v->dx / 2.0 - sw / 2.0;
          const double py = j * v->dy + v->dy / 2.0 - sh / 2.0;
          if (sqrt(px * px + py * py) >= cut - 5.0 * fw)
            img[ORC_AT(i, j, k)] = keep[ORC_AT(i, j, k)];
        }
  }
#undef ORC_AT
  free(keep);
  free(t1);
  free(t2);
}

/* One OSEM sub-iteration's projector pass (reconAlgos.cc:119-159): for each
 * LOR of (subset): l = A x (+ bias); if l > eps, bp += A^T (y / l). */
static void ORC_SFX(osem_pass)(const orc_grid* g, const orc_proj* p,
                               const double* xy, const double* z,
                               const ORC_REAL* x, const ORC_REAL* y,
                               const ORC_REAL* bias, ORC_REAL* bp,
                               int n_subsets, int subset)
{
  const int vps = p->n_views / n_subsets;
  for (int seg = -p->seg_off; seg <= p->seg_off; ++seg)
  {
    const int na = p->n_axial[seg + p->seg_off];
    const long long base = p->seg_base[seg + p->seg_off];
    const long long nlor = (long long)vps * na * p->n_tang;
#pragma omp parallel
    {
      int* ti = (int*)malloc(sizeof(int) * (g->max_path + 16));
      double* tl = (double*)malloc(sizeof(double) * (g->max_path + 16));
#pragma omp for schedule(static)
      for (long long index = 0; index < nlor; ++index)
      {
        int view, axial, tang;
        long long bin;
        double e1[3], e2[3];
        subset_decode(p, na, n_subsets, subset, index, &view, &axial, &tang,
                      &bin);
        lor_endpoints(p, xy, z, seg, view, axial, tang, e1, e2);
        const int n = grid_path(g, e1, e2, ti, tl, g->max_path + 16);
        ORC_REAL line = 0;
        for (int e = 0; e < n; ++e) line += tl[e] * x[ti[e]];
        if (bias) line += bias[base + bin];
        if (line > ORC_EPS)
        {
          const ORC_REAL ratio = y[base + bin] / line;
          for (int e = 0; e < n; ++e)
          {
#pragma omp atomic
            bp[ti[e]] += tl[e] * ratio;
          }
        }
      }
      free(ti);
      free(tl);
    }
  }
}

/* One complete sub-iteration on x (reconAlgos.cc:119-186): projector pass,
 * bp /= sens_s, x *= bp, optional Gaussian, circular cut. bp must be zero
 * on entry; it is left holding the ratio image (not reset). */
void ORC_SFX(orc_osem_subiter)(const orc_vol* v, int geometric,
                               const orc_proj* p, const double* xy,
                               const double* z, const ORC_REAL* y,
                               const ORC_REAL* bias, const ORC_REAL* sens_s,
                               ORC_REAL* x, ORC_REAL* bp, int n_subsets,
                               int subset, int do_convolve, const float* fwhm,
                               float cut, int nthreads)
{
  orc_grid g;
  grid_init(&g, v, geometric);
  set_threads(nthreads);
  const long long nvox = (long long)v->nx * v->ny * v->nz;
  ORC_SFX(osem_pass)(&g, p, xy, z, x, y, bias, bp, n_subsets, subset);
  ORC_SFX(orc_div_clamped)(bp, sens_s, nvox);
  ORC_SFX(orc_mul_clamped)(x, bp, nvox);
  if (do_convolve) ORC_SFX(orc_convolve)(v, x, fwhm, cut);
  ORC_SFX(orc_cut_circle)(v, x, cut);
}

/* Full OSEM (reconAlgos.cc:63-201) on x in place. sens has n_subsets
 * frames; bias may be NULL. Returns 0, or 1 if n_subsets does not divide
 * the number of views (ProjData.cc:166-174). */
int ORC_SFX(orc_osem)(const orc_vol* v, int geometric, const orc_proj* p,
                      const double* xy, const double* z, const ORC_REAL* y,
                      const ORC_REAL* bias, const ORC_REAL* sens, ORC_REAL* x,
                      int n_iter, int n_subsets, float cut, int conv_interval,
                      const float* fwhm, int nthreads)
{
  if (n_subsets <= 0 || p->n_views % n_subsets != 0) return 1;
  const long long nvox = (long long)v->nx * v->ny * v->nz;
  const int conv = conv_interval > 0 && fwhm[0] > 0.0 && fwhm[1] > 0.0 &&
    fwhm[2] > 0.0;
  ORC_REAL* bp = (ORC_REAL*)calloc(nvox, sizeof(ORC_REAL));
  ORC_SFX(orc_cut_circle)(v, x, cut);
  const int total = n_iter * n_subsets;
  for (int it = 0; it < n_iter; ++it)
    for (int s = 0; s < n_subsets; ++s)
    {
      const int sub = it * n_subsets + s + 1;
      ORC_SFX(orc_osem_subiter)(v, geometric, p, xy, z, y, bias,
                                sens + nvox * s, x, bp, n_subsets, s,
                                conv && sub % conv_interval == 0, fwhm, cut,
                                nthreads);
      if (sub != total) memset(bp, 0, sizeof(ORC_REAL) * nvox);
    }
  free(bp);
  return 0;
}

/* Resolution-modelled OSEM as the reference computes it
 * (reconAlgos.cc:203-367), on x in place: cut; every sensitivity frame
 * convolved once (operations::convolve loops over frames,
 * operations.cc:77-80); per sub-iteration the projector pass of x, bp
 * convolved, x /= sens_s then x *= bp (clamped), bp reset, the Gaussian of
 * x every conv_interval sub-iterations, cut. The reference also convolves
 * a copy of x every sub-iteration that no later step reads
 * (reconAlgos.cc:246-247, 271-274, 349-353); it is left out. sens is
 * modified. Returns as orc_osem. */
int ORC_SFX(orc_osem_reso)(const orc_vol* v, int geometric, const orc_proj* p,
                           const double* xy, const double* z,
                           const ORC_REAL* y, const ORC_REAL* bias,
                           ORC_REAL* sens, ORC_REAL* x, int n_iter,
                           int n_subsets, float cut, int conv_interval,
                           const float* fwhm, int nthreads)
{
  if (n_subsets <= 0 || p->n_views % n_subsets != 0) return 1;
  orc_grid g;
  grid_init(&g, v, geometric);
  set_threads(nthreads);
  const long long nvox = (long long)v->nx * v->ny * v->nz;
  const int conv = conv_interval > 0 && fwhm[0] > 0.0 && fwhm[1] > 0.0 &&
    fwhm[2] > 0.0;
  ORC_REAL* bp = (ORC_REAL*)calloc(nvox, sizeof(ORC_REAL));
  ORC_SFX(orc_cut_circle)(v, x, cut);
  for (int s = 0; s < n_subsets; ++s)
    ORC_SFX(orc_convolve)(v, sens + nvox * s, fwhm, cut);
  const int total = n_iter * n_subsets;
  for (int it = 0; it < n_iter; ++it)
    for (int s = 0; s < n_subsets; ++s)
    {
      const int sub = it * n_subsets + s + 1;
      ORC_SFX(osem_pass)(&g, p, xy, z, x, y, bias, bp, n_subsets, s);
      ORC_SFX(orc_convolve)(v, bp, fwhm, cut);
      ORC_SFX(orc_div_clamped)(x, sens + nvox * s, nvox);
      ORC_SFX(orc_mul_clamped)(x, bp, nvox);
      if (sub != total) memset(bp, 0, sizeof(ORC_REAL) * nvox);
      if (conv && sub % conv_interval == 0)
        ORC_SFX(orc_convolve)(v, x, fwhm, cut);
      ORC_SFX(orc_cut_circle)(v, x, cut);
    }
  free(bp);
  return 0;
}
