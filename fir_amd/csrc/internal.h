// internal.h -- the projector context shared by the translation units of
// libfir_hip.so, and the helpers they share. Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/fir_hip.h"
#include "siddon.h"

struct RowMatrix; // LOR-major path matrix of a path cache (path_cache.hip)

struct fir_ctx
{
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;

  // scanner tables
  double2* d_xy = nullptr;
  double* d_z = nullptr;
  int n_cpr = 0;
  int n_slices = 0;

  // sinogram geometry
  bool proj_set = false;
  int n_rings = 0, span = 1, n_seg = 0, n_tang = 0;
  int seg_off = 0, tang_off = 0, n_views = 0, half_span = 0;
  int64_t n_bins = 0;
  std::vector<int> n_axial;
  std::vector<long long> seg_base;
  std::vector<int2> slices; // slice pair per (segment rank, axial), flattened
  int n_pairs = 0;
  int4* d_pairs = nullptr;  // rows in processing order (see fir_set_lor_order)
  int4* d_pairs_bins = nullptr; // rows in sinogram (segment, axial) order
  int* d_ax_prefix = nullptr;   // offsets of each segment in d_pairs_bins
  int* d_n_axial = nullptr;
  long long* d_seg_base = nullptr;
  int lor_order = FIR_ORDER_SEGMENT;
  int row_mode = fir::ROWS_SEGMENT_VIEW;
  int tang_major = 0;

  // private per-XCD copies of scatter targets (1: scatter directly)
  int bp_copies = 1;
  float* d_copies = nullptr;
  size_t copies_floats = 0;

  int walk_mode = FIR_WALK_REFERENCE;
  // allow subset counts that do not divide the number of views
  bool uneven_subsets = false;

  // 64-bit device counter of fir_path_count
  unsigned long long* d_counter = nullptr;

  // scratch images of the Gaussian (library-allocated, grown on demand)
  float* d_scratch = nullptr;
  size_t scratch_floats = 0;
  // Gaussian taps on the device: [x: 64][y: 64][z: 64]
  // Gaussian taps, one set per (FWHM, voxel size), each uploaded once and
  // kept until the context is destroyed (captured graphs hold their
  // addresses): X taps at [0, 64), Y at [64, 128), Z at [128, 192).
  struct TapSet
  {
    float key[3];
    double vox[3];
    int h[3];
    float* d = nullptr;
  };
  std::vector<TapSet> tap_sets;
  // Scratch and scatter-copy buffers replaced by larger ones: freed with
  // the context, not before, since captured graphs may still use them.
  std::vector<void*> retired;

  // Bumped whenever the scanner, sinogram or walk mode changes: path
  // matrices built under another generation are not shared.
  unsigned generation = 0;
  // Lanes per LOR and per voxel the path caches built on this context are
  // laid out for (fir_set_cache_lanes).
  int cache_lanes[2] = {8, 16};
  // z-slices interleaved in the gather layouts of the path caches built on
  // this context (fir_set_cache_zslices).
  int cache_zq = 4;
  // Largest number of path elements one transpose step sorts (its
  // temporaries take 24 bytes per element).
  long long cache_chunk = 1LL << 28;
  // Whether the path matrices built afterwards get the sheet form.
  int cache_sheets = 1;
  // Whether the walked OSEM projector runs family by family with the
  // family's sheet in LDS (fir_set_walk_sheets).
  int walk_sheets = 1;
  // Path matrices of live caches, shared by caches of one volume, walk mode
  // and length format.
  std::vector<std::weak_ptr<RowMatrix>> row_matrices;
};

namespace firi
{
// Sets the thread's last-error message and returns code.
int fail(int code, const char* fmt, ...);

constexpr int kBlock = 256;

unsigned blocks_for(long long n, int block = kBlock);
// Grid of element-wise kernels (grid-stride loops).
unsigned ew_blocks(long long n);
int make_grid(const fir_ctx* ctx, const fir_vol_desc* v, fir::Grid* g);
int need_lors(const fir_ctx* ctx);
int check_subsets(const fir_ctx* ctx, int n_subsets);
fir::LorGeom lor_geom(const fir_ctx* ctx);
// LorGeom whose pair table is in sinogram (segment, axial) order.
fir::LorGeom lor_geom_bins(const fir_ctx* ctx);
long long n_work(const fir_ctx* ctx, int n_subsets, int subset);
int check_work(long long n);

// One image of a multi-job OSEM step (fir_osem_step_run), for its
// element-wise phases: the update and the Gaussian.
struct StepImage
{
  float* x;
  float* bp;
  const float* sens;
  float* t1;          // the Gaussian's intermediate images (nx*ny*nz each)
  float* t2;
  const float* taps;  // X, Y, Z taps, 64 floats each (NULL: no Gaussian)
  int h[3];
  int nx, ny, nz;
  double dx, dy;
  int tile0;          // the image's first tile (blocks of 32 x 8 voxels of
                      // one slice) in the step's tile numbering
  int vt0[3];         // its first block in each vectorised Gaussian pass
};
// Tiles of one image in that numbering.
int step_image_tiles(int nx, int ny, int nz);
// Blocks of one image in the vectorised Gaussian pass of an axis, and
// whether an image of that width and those half-widths can take them.
int step_image_vtiles(int nx, int ny, int nz, int axis);
bool gauss_vec_shape(int nx, const int* h);
// The context's taps of the Gaussian of fwhm on grid v (uploaded once and
// kept until the context is destroyed).
int gauss_taps_for(fir_ctx* c, const fir_vol_desc* v, const float* fwhm,
                   const float** taps, int h[3]);
// The element-wise phases of a step over n_imgs images (device array),
// n_tiles tiles in all: the OSEM update, then (gauss) the three passes of
// the Gaussian, the Z pass with the edge restore at radius restore_r (when
// restore) and the circular cut; without the Gaussian the update cuts.
// With vtiles (the blocks of each vectorised pass over all images) the
// passes are the vectorised ones, for half-widths vh shared by every image.
int launch_step_elementwise(hipStream_t st, const StepImage* imgs, int n_imgs,
                            int n_tiles, bool gauss, float cut, int restore,
                            double restore_r, const int* vtiles,
                            const int* vh);
} // namespace firi

#define FIR_HIP_CHECK(expr)                                                  \
  do                                                                         \
  {                                                                          \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(FIR_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

#define FIR_CHECK_LAUNCH()                                                   \
  do                                                                         \
  {                                                                          \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess)                                                    \
      return fail(FIR_E_HIP, "kernel launch: %s", hipGetErrorString(e_));    \
  } while (0)
