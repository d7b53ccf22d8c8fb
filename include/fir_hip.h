/*
 * fir_hip.h -- C-ABI of the MI355X (gfx950) implementation of FIR's
 * projector and OSEM hot path.
 *
 * Conventions
 *   - Every function returns 0 (FIR_OK) or a nonzero FIR_E_* code; the
 *     message of the last failure on the calling thread is returned by
 *     fir_last_error(). Nothing throws across this boundary. This mirrors the
 *     reference's error() exceptions (console.inl:74-82), which the C++ host
 *     layer re-raises from these codes.
 *   - Image, sinogram, sensitivity and scratch pointers are DEVICE pointers
 *     (hipMalloc'd, or fir_malloc'd). Host pointers appear only where named
 *     (geometry tables and the h2d/d2h copies).
 *   - Work is enqueued asynchronously on the context's stream
 *     (fir_set_stream); fir_synchronize() waits for it. Calls on one
 *     context must not overlap; different contexts may run concurrently.
 *   - HIP graph capture: projector, update, Gaussian, cut and path-cache
 *     calls may be captured once the same call has run eagerly on the
 *     context (which uploads the Gaussian taps of that FWHM and voxel size
 *     and sizes the scratch). Taps are kept per (FWHM, voxel size), and
 *     scratch buffers replaced by larger ones stay allocated until the
 *     context is destroyed, so later calls never invalidate a captured
 *     graph.
 *   - Images are FP32, x fastest, then y, then z (reference VolData.inl:53-56),
 *     frames contiguous. Sinograms are FP32 in the concatenated segment order
 *     of the reference .s file: segment -S..S, each [view][axial][tang]
 *     (reference ProjData.h:9-21, ProjInterfileReader.cc:79-93).
 */
#ifndef FIR_HIP_H
#define FIR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum
{
  FIR_OK = 0,
  FIR_E_ARG = 1,   /* invalid argument or geometry */
  FIR_E_HIP = 2,   /* HIP runtime failure */
  FIR_E_STATE = 3, /* call made before the geometry it needs was set */
  FIR_E_NOMEM = 4  /* device allocation failed */
};

/* Siddon step extents: the reference truncates voxel extents to int for the
 * walk (Siddon.h:15-16,113); FIR_WALK_GEOMETRIC uses the true extents. */
enum
{
  FIR_WALK_REFERENCE = 0,
  FIR_WALK_GEOMETRIC = 1
};

/* Order in which the projector kernels visit the rows of a sinogram
 * (results do not depend on it beyond FP32 summation order). The
 * (segment, axial) pairs are ranked by segment then axial
 * (FIR_ORDER_SEGMENT) or by the centre slice of their LORs
 * (FIR_ORDER_ZCENTER). A subset's rows are visited by default segment by
 * segment, each segment view by view and each view axial position by axial
 * position (the sinogram's own order; needs FIR_ORDER_SEGMENT): concurrent
 * LORs then spread over many z-slabs and few views of one ring difference,
 * which keeps backprojection atomics on distinct voxels.
 * FIR_ORDER_VIEW_MAJOR visits every pair of one view together, and
 * FIR_ORDER_PAIR_MAJOR every view of one pair. With FIR_ORDER_TANG_MAJOR,
 * neighbouring lanes take the same tangential bin of consecutive rows
 * instead of consecutive bins of one row. */
enum
{
  FIR_ORDER_SEGMENT = 0,
  FIR_ORDER_ZCENTER = 1,
  FIR_ORDER_PAIR_MAJOR = 0x100,
  FIR_ORDER_TANG_MAJOR = 0x200,
  FIR_ORDER_VIEW_MAJOR = 0x400
};

typedef struct fir_ctx fir_ctx;

/* Volume geometry: the reference VolHeader (VolHeader.h:6-34). */
typedef struct
{
  int nx, ny, nz;    /* voxels per axis */
  double dx, dy, dz; /* voxel extent, mm */
  double ox, oy, oz; /* x,y: centre of the first voxel; z: volume-centre offset */
} fir_vol_desc;

/* Message of the last failed call on this thread ("" if none). */
const char* fir_last_error(void);

/* ABI version of this library (incremented on incompatible changes). */
int fir_abi_version(void);

/* ---- context -------------------------------------------------------- */

/* Create a context on HIP device `device` with its own stream. */
int fir_ctx_create(fir_ctx** ctx, int device);
int fir_ctx_destroy(fir_ctx* ctx);
/* Enqueue on an external hipStream_t (NULL: the device's null stream).
 * A new context uses a non-blocking stream of its own. */
int fir_set_stream(fir_ctx* ctx, void* hip_stream);
int fir_synchronize(fir_ctx* ctx);

/* ---- geometry ------------------------------------------------------- */

/* Scanner position tables, computed by the host as the reference
 * ScannerData does (ScannerData.cc:248-377): crystal_xy[2*n_cpr] holds (x,y)
 * of each angular crystal index, slice_z[n_slices] the z of each slice
 * (n_slices = 2*n_rings - 1). Host pointers; copied to the device. */
int fir_set_scanner(fir_ctx* ctx, const double* crystal_xy, int n_cpr,
                    const double* slice_z, int n_slices);

/* Sinogram geometry from the reference ProjHeader keys (ProjHeader.h:15-40);
 * validated as ProjHeader::check (ProjHeader.cc:6-106) and derived as
 * ProjGeometry::fill (ProjHeader.cc:108-186). n_tang = 0 selects n_cpr - 1. */
int fir_set_proj(fir_ctx* ctx, int n_rings, int n_cpr, int span,
                 int n_segments, int n_tang);

/* Number of bins of the current sinogram geometry (0 if unset). */
int64_t fir_proj_n_bins(const fir_ctx* ctx);

/* FIR_WALK_REFERENCE (default) or FIR_WALK_GEOMETRIC. */
int fir_set_walk_mode(fir_ctx* ctx, int mode);

/* Subset counts: by default n_subsets must divide the number of views, as
 * the reference requires (ProjData.cc:166-174). With allow != 0 any count
 * from 1 to nViews is accepted; subset s is still the views
 * v = k * n_subsets + s, so the first nViews % n_subsets subsets hold one
 * view more (an extension: the reference defines no such split). */
int fir_set_uneven_subsets(fir_ctx* ctx, int allow);

/* Kernel tuning (no effect on results beyond FP32 summation order):
 * the row order (FIR_ORDER_*), and the number of private copies (1, 2, 4
 * or 8) the scatter kernels accumulate into, one per XCD, summed into the
 * target afterwards. Copies cost copies x N_vox floats of device memory
 * (allocated by the library on first use). */
int fir_set_lor_order(fir_ctx* ctx, int order);
int fir_set_scatter_copies(fir_ctx* ctx, int copies);

/* Whether fir_osem_project (the walked OSEM projector) runs family by
 * family (default 1, when the image's sheets fit in 64 KB of LDS): the
 * LORs of one (view, tangential bin) cross the same pixels, so one block
 * stages that sheet of x (pixels x slices) in LDS for the line integrals
 * (the same sums bit for bit), then accumulates the family's
 * backprojection there and adds each touched voxel to bp once. 0 walks
 * LOR by LOR with one atomic per path element. */
int fir_set_walk_sheets(fir_ctx* ctx, int enable);

/* ---- device memory -------------------------------------------------- */

int fir_malloc(fir_ctx* ctx, void** dptr, size_t bytes);
int fir_free(fir_ctx* ctx, void* dptr);
int fir_memcpy_h2d(fir_ctx* ctx, void* dst, const void* src, size_t bytes);
int fir_memcpy_d2h(fir_ctx* ctx, void* dst, const void* src, size_t bytes);
/* The same copies enqueued on the context's stream without waiting for
 * them (fir_synchronize waits). The host memory must stay valid and
 * unchanged (h2d) or unread (d2h) until then; from page-locked memory
 * (fir_host_alloc) the copy runs as DMA while the host goes on, e.g.
 * reading the next chunk of a file. No counterpart in the reference. */
int fir_memcpy_h2d_async(fir_ctx* ctx, void* dst, const void* src,
                         size_t bytes);
int fir_memcpy_d2h_async(fir_ctx* ctx, void* dst, const void* src,
                         size_t bytes);
/* Page-locked host memory (hipHostMalloc): copies to and from it run as
 * direct DMA instead of through a pageable staging copy. No counterpart in
 * the reference (its containers malloc their data, ProjData.cc:623-651,
 * VolData.cc:393-418); the host containers keep sinograms and volumes in
 * it. FIR_E_NOMEM if it cannot be allocated (for instance with no device
 * present). */
int fir_host_alloc(void** ptr, size_t bytes);
int fir_host_free(void* ptr);
/* Fill n floats with value (a kernel, safe under graph capture). */
int fir_fill(fir_ctx* ctx, float* dst, int64_t n, float value);

/* ---- projectors ----------------------------------------------------- */

/* Forward projection of every bin: replaces projections::forward
 * (projections.cc:15-92). sino gets fir_proj_n_bins() floats; bins whose
 * LOR misses the volume get 0. */
int fir_forward(fir_ctx* ctx, const fir_vol_desc* vol, const float* img,
                float* sino);

/* Forward projection of the bins of one subset (views v with
 * v % n_subsets == subset, reference LORCache.cc:49-51); other bins are
 * left untouched. */
int fir_forward_subset(fir_ctx* ctx, const fir_vol_desc* vol,
                       const float* img, float* sino, int subset,
                       int n_subsets);

/* Backprojection: replaces projections::backward (projections.cc:94-184).
 * img has n_subsets frames; all are zeroed, then frame s receives the
 * scatter of subset s. */
int fir_backward(fir_ctx* ctx, const fir_vol_desc* vol, const float* sino,
                 float* img, int n_subsets);

/* Backprojection of one subset, accumulated into img (not zeroed). A NULL
 * sino backprojects ones. */
int fir_backward_subset(fir_ctx* ctx, const fir_vol_desc* vol,
                        const float* sino, float* img, int subset,
                        int n_subsets);

/* Sensitivity: replaces projections::computeSensitivityVol
 * (projections.cc:186-197): backprojection of ones, one frame per subset. */
int fir_sensitivity(fir_ctx* ctx, const fir_vol_desc* vol, float* sens,
                    int n_subsets);

/* ---- OSEM ----------------------------------------------------------- */

/* Projector pass of one OSEM sub-iteration (reconAlgos.cc:119-159): for
 * every LOR of the subset, l = A x (+ bias); if l > 1e-5,
 * bp += A^T (y / l). bias may be NULL. */
int fir_osem_project(fir_ctx* ctx, const fir_vol_desc* vol, const float* x,
                     const float* y, const float* bias, float* bp, int subset,
                     int n_subsets);

/* Projector pass for n_images reconstructions that share the image
 * geometry and the subset schedule (noise realisations, dynamic frames):
 * x, y, bias, bp are HOST arrays of n_images DEVICE pointers (bias may be
 * NULL, or hold NULL entries). Each LOR is walked once for up to 8 images;
 * each image gets exactly what fir_osem_project would give it. */
int fir_osem_project_batch(fir_ctx* ctx, const fir_vol_desc* vol,
                           int n_images, const float* const* x,
                           const float* const* y, const float* const* bias,
                           float* const* bp, int subset, int n_subsets);

/* Update of one sub-iteration (reconAlgos.cc:161-186, VolData.cc:305-343):
 * bp = clamp(bp / sens_s), x = clamp(x * bp), then the circular cut of x
 * when cut_radius > 0 and no convolution follows, and bp = 0 when
 * reset_bp != 0. */
int fir_osem_update(fir_ctx* ctx, const fir_vol_desc* vol, float* x,
                    float* bp, const float* sens_s, float cut_radius,
                    int apply_cut, int reset_bp);

/* One complete sub-iteration: project, update, optional Gaussian
 * (do_convolve != 0) with edge restore, circular cut, bp reset. bp must be
 * zero on entry. */
int fir_osem_subiter(fir_ctx* ctx, const fir_vol_desc* vol, float* x,
                     const float* y, const float* bias, const float* sens_s,
                     float* bp, int subset, int n_subsets, int do_convolve,
                     const float* fwhm_xyz, float cut_radius, int reset_bp);

/* ---- differentiable projector --------------------------------------- */

/* Poisson negative log-likelihood of x for data y with additive term r:
 * *loss = sum_b (l_b - y_b log l_b), l = A x + r (r may be NULL), and
 * grad = A^T (1 - y / l), the exact adjoint of fir_forward applied to the
 * weights. Bins with l <= 0 contribute l and weight 1. grad (N_vox floats)
 * and loss (one double) are device pointers, both overwritten. No
 * counterpart in the reference (its OSEM needs no gradient); it backs the
 * PyTorch autograd function of the projector. */
int fir_poisson_nll_grad(fir_ctx* ctx, const fir_vol_desc* vol,
                         const float* x, const float* y, const float* r,
                         float* grad, double* loss);

/* ---- image / sinogram operations ------------------------------------ */

/* Separable Gaussian with border renormalisation and edge restore:
 * replaces operations::convolve (operations.cc:11-192) on n_frames frames.
 * fwhm_xyz is a host pointer to 3 floats. */
int fir_convolve(fir_ctx* ctx, const fir_vol_desc* vol, float* img,
                 int n_frames, const float* fwhm_xyz, float cut_radius);

/* fir_convolve without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Same kernels and the
 * same result bit for bit as fir_convolve; its taps and two intermediate
 * images live in a static device buffer that the library holds on every
 * device it runs on (about 64 MB, until the process ends) instead of the
 * context's scratch, and the taps travel as kernel arguments. Frames of
 * at most 8 388 608 voxels (2^23, e.g. 256x256x128); larger ones fail
 * with FIR_E_ARG. All calls on one device share that buffer: calls that
 * may overlap on one device (two streams, two contexts, two threads) race
 * and must be serialised by the caller. */
int fir_convolve_noalloc(fir_ctx* ctx, const fir_vol_desc* vol, float* img,
                         int n_frames, const float* fwhm_xyz,
                         float cut_radius);

/* operations::cutCircle (operations.cc:194-220) on n_frames frames. */
int fir_cut_circle(fir_ctx* ctx, const fir_vol_desc* vol, float* img,
                   int n_frames, float cut_radius);

/* a = (a > 1e-5 && b > 1e-5) ? a*b : 0 -- VolData::operator*=
 * (VolData.cc:305-323) and ProjData::operator*= (ProjData.cc:481-512). */
int fir_mul_clamped(fir_ctx* ctx, float* a, const float* b, int64_t n);

/* a = (a > 1e-5 && b > 1e-5) ? a/b : 0 -- VolData::operator/=
 * (VolData.cc:325-343). */
int fir_div_clamped(fir_ctx* ctx, float* a, const float* b, int64_t n);

/* a = a > 1e-5 ? exp(a) : 1 -- ProjData::exponential (ProjData.cc:514-539). */
int fir_exp_clamped(fir_ctx* ctx, float* a, int64_t n);

/* Hounsfield units to mu (1/mm) -- operations::HounsfieldToMuMap
 * (operations.cc:243-284). */
int fir_hu_to_mu(fir_ctx* ctx, float* img, int64_t n);

/* ---- parity helper -------------------------------------------------- */

/* Siddon paths of bins [bin0, bin0 + n_bins) (all-bins order): counts[b]
 * = number of elements (-1 if the LOR misses the volume); idx/len receive
 * the elements of bin b at [b * max_per_lor, ...). Replaces the path
 * computed by Siddon::computePath (Siddon.cc:112-258). */
int fir_path_dump(fir_ctx* ctx, const fir_vol_desc* vol, int64_t bin0,
                  int64_t n_bins, int max_per_lor, int* counts, int* idx,
                  double* len);

/* Total number of Siddon path elements (reference semantics) over the LORs
 * of one subset (n_subsets = 0: every bin); *total is a host pointer. Used
 * for the algorithmic-bytes count of the projector kernels. Synchronizes
 * the context's stream. */
int fir_path_count(fir_ctx* ctx, const fir_vol_desc* vol, int subset,
                   int n_subsets, int64_t* total);

/* ---- list-mode events and span weighting ----------------------------- */

/* List-mode events into sinogram counts: ProjData::getBinCoordinates
 * (ProjData.cc:342-479) followed by ProjData::incrementBin
 * (ProjData.inl:47-56), for every event at once. events: n_events x 4
 * int32 on the device (ring1, crystal1, ring2, crystal2), 16-byte aligned.
 * sino[bin] += weights ? weights[i] : 1 for every event whose crystal pair
 * the sinogram holds; events with rings or crystals out of range, a ring
 * difference beyond the last segment or a tangential coordinate outside
 * the stored range are skipped. sino is not cleared. If n_counted is not
 * NULL it receives the number of events counted (the call then waits for
 * the context's stream). */
int fir_histogram_events(fir_ctx* ctx, const int* events, int64_t n_events,
                         const float* weights, float* sino,
                         int64_t* n_counted);

/* ProjData::rebinWeight (ProjData.cc:541-615) on a device sinogram: with a
 * segment span above 1, every bin of a (segment, axial) position that fuses
 * k > 1 ring pairs is multiplied by (float)(1.0 / k). No-op for span 1. */
int fir_rebin_weight(fir_ctx* ctx, float* sino);

/* ---- path cache ------------------------------------------------------ */

/* Every LOR's Siddon path, walked once and kept in device memory: the
 * LOR-major path matrix (shared by the caches of one context, volume, walk
 * mode and length format) plus, per cache, its voxel-major transpose for
 * one subset split. Projector passes then stream the matrix: forward
 * projection and OSEM line integrals gather over a LOR's elements,
 * backprojection gathers over a voxel's elements (no atomics; results do
 * not depend on scheduling). It extends the reference's LORCache
 * (LORCache.h:5-47, LORCache.cc:10-168), which keeps the crystal pairs of
 * each (subset, segment), to whole paths. A cache describes the geometry
 * at creation; later fir_set_scanner / fir_set_proj / fir_set_walk_mode
 * calls do not change it. Calls on one cache must not overlap (they share
 * its per-LOR weight buffer). */
typedef struct fir_cache fir_cache;

/* Stored element length: FIR_LEN_SNORM16 (default) round(32767 * len / m)
 * with m the LOR's largest |length| (the reference walk with int-truncated
 * steps, FIR_WALK_REFERENCE on non-integer voxel sizes, yields some negative
 * lengths), FIR_LEN_FP16 binary16 of len / m, FIR_LEN_FP32 binary32 of len;
 * each restored with a per-LOR FP32 factor. */
enum
{
  FIR_LEN_SNORM16 = 0,
  FIR_LEN_FP16 = 1,
  FIR_LEN_FP32 = 2
};

/* Lanes per LOR (forward) and per voxel (backprojection) that the caches
 * created afterwards on this context are laid out for, powers of two from 1
 * to 64 (default 8 and 16); also their initial fir_cache_tune setting.
 * Within every path and voxel list, each round of `lanes` quads is stored
 * so that the lanes' loads read consecutive elements. */
int fir_set_cache_lanes(fir_ctx* ctx, int fwd_lanes, int bp_lanes);

/* z-slices interleaved innermost in the image layouts the caches created
 * afterwards on this context gather from: 1, 2, 4, 8 or 16 (default 4).
 * The LORs a wave projects together cross one line of the xy-plane at
 * neighbouring centre slices; interleaving z lets their voxels share cache
 * lines (one image: -8% per projector pass at 4; with 8 interleaved images
 * 1 or 2 is faster). Results do not depend on it. */
int fir_set_cache_zslices(fir_ctx* ctx, int zslices);

/* Largest number of path elements one step of a cache build's transpose
 * sorts, 1024 to 2^30 (default 2^28); the step's temporaries take 24 bytes
 * per element, so smaller steps fit beside larger caches. A subset's
 * voxel lists come out the same whatever the step. */
int fir_set_cache_chunk(fir_ctx* ctx, int64_t max_elements);

/* Whether the path matrices built afterwards on this context also get the
 * sheet form (default 1): with rows in tangential-major order, the LORs of
 * one (view, tangential bin) cross the same pixels, and each path element
 * gets a 16-bit index within that family's sheet (its pixels x all
 * slices), stored with its 16-bit length in one word, so that one-image
 * OSEM passes stage the sheet in LDS and gather there (fir_cache_tune,
 * FIR_CACHE_NO_SHEETS). 16-bit length formats only; it costs 4 bytes per
 * element (15.6 GB for 3.9e9 elements); 0 saves that memory. */
int fir_set_cache_sheets(fir_ctx* ctx, int enable);

/* Build the cache of `vol` for `n_subsets` subsets (views v with
 * v % n_subsets == s, as LORCache.cc:49-51). Synchronizes the context's
 * stream; FIR_E_NOMEM if the matrices do not fit in device memory. */
int fir_cache_create(fir_ctx* ctx, const fir_vol_desc* vol, int n_subsets,
                     int len_format, fir_cache** out);

/* A cache of subsets 0 .. k-1 of n_subsets only, for scanners whose whole
 * matrix does not fit in device memory (the reference's LORCache holds
 * every subset, LORCache.cc:10-96): k is the largest count, at most
 * max_subsets (<= 0: no cap), whose matrices and build temporaries fit in
 * max_bytes (<= 0: no limit) and in the free device memory; FIR_E_NOMEM
 * if subset 0 alone does not. The other subsets are projected with the
 * on-the-fly walk (fir_osem_project, fir_forward, fir_backward): on a
 * partial cache fir_cache_osem_project of such a subset fails with
 * FIR_E_ARG, fir_cache_forward writes the bins of the cached subsets'
 * views only, and fir_cache_backward / fir_cache_sensitivity write the
 * frames of the cached subsets only (other bins and frames are left as
 * they are). Its path matrix is its own, not shared. */
int fir_cache_create_partial(fir_ctx* ctx, const fir_vol_desc* vol,
                             int n_subsets, int len_format, int max_subsets,
                             int64_t max_bytes, fir_cache** out);

/* Subset count of the cache and how many (0 .. n_cached-1) it holds. */
int fir_cache_subsets(const fir_cache* cache, int* n_subsets, int* n_cached);

int fir_cache_destroy(fir_cache* cache);

/* Crossing LORs, path elements, and device bytes held by the cache
 * (including its share of the path matrix). Any pointer may be NULL. */
int fir_cache_info(const fir_cache* cache, int64_t* n_rows,
                   int64_t* n_elements, int64_t* device_bytes);

/* Which forms of the path matrix the cache holds, as bits of *forms:
 * FIR_FORM_SHEETS, the sheet form (16-bit sheet indices of tangential-major
 * rows), and FIR_FORM_FAMILY_LISTS, the family lists plus the partial
 * sheets of the family backprojection. The build adds each only when the
 * device has room for it, and the lists only when the rows of consecutive
 * elements of each voxel's list lie at most 127 apart (their stored quads
 * hold row steps in 7 bits); otherwise one-image passes fall back to the
 * image gathers and the transposed lists, with results within FP32
 * rounding; one-image OSEM passes use what is held unless fir_cache_tune
 * switches it off. FIR_FORM_SLABS: batches of 4 or 8 images take their
 * line integrals from sheets staged slab by slab (sheet pixels ranked along
 * each family's line, checked at build; the runtime forward lanes equal to
 * the layout's). No counterpart in the reference. */
enum
{
  FIR_FORM_SHEETS = 1,
  FIR_FORM_FAMILY_LISTS = 2,
  FIR_FORM_SLABS = 4
};
int fir_cache_forms(const fir_cache* cache, int* forms);

/* Kernel tuning (results change only by FP32 summation order): lanes
 * that share one LOR (forward) and one voxel (backprojection), powers of
 * two from 1 to 64 (default: the layout's, fir_set_cache_lanes; other
 * values give the same results, with less locality). The path matrix's
 * row order within a view follows the context's row order
 * (fir_set_lor_order: pair order, FIR_ORDER_TANG_MAJOR) at creation.
 * Backprojections of one or two images walk the voxels z fastest (a wave's
 * voxels are then crossed by neighbouring rows). Flags: FIR_CACHE_XCD_MAP
 * deals contiguous eighths of the rows of a view (of the voxels) to the
 * eight XCDs; FIR_CACHE_STORED_VOXELS walks the voxels in storage order.
 * One-image OSEM passes of a cache whose rows are in tangential-major order
 * (FIR_ORDER_VIEW_MAJOR | FIR_ORDER_TANG_MAJOR) compute their line
 * integrals family by family: the LORs of one (view, tangential bin) cross
 * the same pixels of the xy-plane, so that sheet of the image (pixels x
 * slices, up to 64 KB) is staged in LDS and the paths index it with 16-bit
 * entries (the same sums bit for bit); FIR_CACHE_NO_SHEETS gathers from
 * the image instead. Such passes also backproject family by family when
 * the matrix holds family lists (built with the sheet form when the device
 * has room: 4 bytes per element plus 8 per sheet entry): every voxel of a
 * family's sheet has the list of the family's elements through it, which
 * the block sums against the family's weights, kept in LDS, into a partial
 * sheet; a second kernel sums, for every voxel, the partials of the
 * subset's families that hold its pixel, in family order. The sums are the
 * same on every run. FIR_CACHE_NO_FAMILY_BP gathers the weights through
 * the transposed voxel lists instead (sums in another order: results
 * within FP32 rounding of each other). Batches of 4 or 8 images with the
 * family lists stage each family's sheet of all their images in LDS slab
 * by slab (pixels along the family's line, all slices) and take their line
 * integrals from it, each image's the same bit for bit as a one-image
 * pass's; FIR_CACHE_NO_SLABS gathers them from interleaved image copies
 * instead (also the same bits). */
enum
{
  FIR_CACHE_XCD_MAP = 1,
  FIR_CACHE_STORED_VOXELS = 2,
  FIR_CACHE_NO_SHEETS = 4,
  FIR_CACHE_NO_FAMILY_BP = 8,
  FIR_CACHE_NO_SLABS = 16
};
int fir_cache_tune(fir_cache* cache, int fwd_lanes, int bp_lanes, int flags);

/* projections::forward (projections.cc:15-92) over the cached paths: every
 * bin written, LORs that miss the volume give 0. */
int fir_cache_forward(fir_cache* cache, const float* img, float* sino);

/* projections::backward (projections.cc:94-184) with the cache's subset
 * count: frames[s * n_voxels ...] = backprojection of subset s (overwritten). */
int fir_cache_backward(fir_cache* cache, const float* sino, float* frames);

/* projections::computeSensitivityVol (projections.cc:186-197): backward of
 * ones, one frame per subset. */
int fir_cache_sensitivity(fir_cache* cache, float* frames);

/* OSEM projector pass of subset `subset` for n_images images
 * (reconAlgos.cc:126-158): bp[b] = A_s^T (y_b / (A_s x_b + bias_b)) over
 * the LORs whose line exceeds 1e-5 (bp[b] is overwritten). Up to 8 images
 * share each read of the matrix. bias may be NULL, or hold NULL entries. */
int fir_cache_osem_project(fir_cache* cache, int n_images,
                           const float* const* x, const float* const* y,
                           const float* const* bias, float* const* bp,
                           int subset);

/* ---- Many OSEM solves per launch (no counterpart in the reference) ----
 * The reference runs one reconstruction at a time (reconAlgos.cc:63-201).
 * An OSEM step here advances many reconstructions by one sub-iteration
 * each: jobs of up to 8 images sharing a path cache, on any number of
 * caches of one context (different image grids and subset counts), every
 * phase of all jobs in one launch. The projector phases are persistent
 * kernels: a grid that fills the device walks one table of (job, family)
 * work items, the most work first, so the launch ends on the smallest
 * families of whichever jobs hold them. */
#define FIR_JOB_MAX_IMAGES 8
typedef struct fir_osem_job
{
  fir_cache* cache;    /* path cache with the sheet form and family lists */
  int subset;          /* the sub-iteration's subset (a held subset) */
  int n_images;        /* 1 .. FIR_JOB_MAX_IMAGES */
  float* x[FIR_JOB_MAX_IMAGES];           /* images, updated in place */
  const float* y[FIR_JOB_MAX_IMAGES];     /* measured sinograms */
  const float* bias[FIR_JOB_MAX_IMAGES];  /* additive terms, or NULL */
  float* bp[FIR_JOB_MAX_IMAGES];          /* per-image scratch images */
  const float* sens;   /* the subset's sensitivity frame (all images) */
  void* workspace;     /* fir_cache_job_workspace bytes, 256-byte aligned;
                          jobs of one step need distinct workspaces */
} fir_osem_job;

/* Device bytes of a job workspace on this cache: per-row weights, the
 * images' pixel-major copies, the family partial sheets and the Gaussian's
 * intermediate images of up to 8 images. */
int fir_cache_job_workspace(const fir_cache* cache, int64_t* bytes);

/* A prepared step: the jobs' tables are uploaded once, so that running it
 * allocates and copies nothing (it can be captured in a HIP graph and
 * replayed). fwhm (NULL or all 0: none) and cut as fir_osem_subiter: for
 * every image, bp = A_s^T (y / (A_s x + bias)) (reconAlgos.cc:126-158),
 * x = x * bp / sens with the clamps of reconAlgos.cc:160-166, then the
 * Gaussian with its edge restore (operations.cc:11-185) and the circular
 * cut (operations.cc:194-220); without a Gaussian, the cut only. bp is
 * left at 0. Every image gets exactly what fir_cache_osem_project,
 * fir_osem_update and fir_convolve + fir_cut_circle give it. All caches
 * must belong to one context, hold the sheet form, family lists and slab
 * form (fir_cache_forms) in one length format, laid out for the same lanes
 * per LOR (fir_set_cache_lanes; 4 suit the steps best). */
typedef struct fir_osem_step fir_osem_step;
int fir_osem_step_create(int n_jobs, const fir_osem_job* jobs,
                         const float* fwhm, float cut, fir_osem_step** out);
/* Runs the step on the context's stream: one launch per phase for all
 * jobs (pack; line integrals and family backprojection, each once for the
 * jobs of 8-image slots and once for those of 4; per-voxel sums; update;
 * the Gaussian's three passes). */
int fir_osem_step_run(fir_osem_step* step);
int fir_osem_step_destroy(fir_osem_step* step);

#ifdef __cplusplus
}
#endif

#endif /* FIR_HIP_H */
