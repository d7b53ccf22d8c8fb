"""ctypes binding of libfir_hip.so (the C-ABI of include/fir_hip.h).

PyTorch is used only for device memory and streams: every operator takes
torch CUDA (HIP) tensors and passes their device pointers through the C-ABI.
There is no fallback: if the HIP library is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfir_hip.so")

FIR_WALK_REFERENCE = 0
FIR_WALK_GEOMETRIC = 1
FIR_ORDER_SEGMENT = 0
FIR_ORDER_ZCENTER = 1
FIR_ORDER_PAIR_MAJOR = 0x100
FIR_ORDER_TANG_MAJOR = 0x200
FIR_ORDER_VIEW_MAJOR = 0x400
FIR_LEN_SNORM16 = 0
FIR_LEN_FP16 = 1
FIR_LEN_FP32 = 2
FIR_CACHE_XCD_MAP = 1
FIR_CACHE_STORED_VOXELS = 2
FIR_CACHE_NO_SHEETS = 4
FIR_CACHE_NO_FAMILY_BP = 8
FIR_CACHE_NO_SLABS = 16
FIR_FORM_SHEETS = 1
FIR_FORM_FAMILY_LISTS = 2
FIR_FORM_SLABS = 4
# Row order of path caches: per view, tangential bin then (segment, axial)
# pair by centre slice -- the rows through one voxel are then close together.
CACHE_ROW_ORDER = FIR_ORDER_VIEW_MAJOR | FIR_ORDER_TANG_MAJOR | FIR_ORDER_ZCENTER

# Every symbol declared in include/fir_hip.h.
EXPORTS = (
    "fir_last_error", "fir_abi_version", "fir_ctx_create", "fir_ctx_destroy",
    "fir_set_stream", "fir_synchronize", "fir_set_scanner", "fir_set_proj",
    "fir_proj_n_bins", "fir_set_walk_mode", "fir_malloc", "fir_free",
    "fir_memcpy_h2d", "fir_memcpy_d2h", "fir_memcpy_h2d_async",
    "fir_memcpy_d2h_async", "fir_host_alloc", "fir_host_free",
    "fir_fill", "fir_forward",
    "fir_forward_subset", "fir_backward", "fir_backward_subset",
    "fir_sensitivity", "fir_osem_project", "fir_osem_update",
    "fir_osem_subiter", "fir_convolve", "fir_convolve_noalloc",
    "fir_cut_circle", "fir_mul_clamped",
    "fir_div_clamped", "fir_exp_clamped", "fir_hu_to_mu", "fir_path_dump",
    "fir_path_count", "fir_poisson_nll_grad", "fir_osem_project_batch",
    "fir_set_uneven_subsets", "fir_set_lor_order", "fir_set_scatter_copies",
    "fir_set_cache_lanes", "fir_set_cache_zslices", "fir_set_cache_chunk",
    "fir_set_cache_sheets", "fir_set_walk_sheets",
    "fir_cache_create",
    "fir_cache_create_partial", "fir_cache_subsets", "fir_cache_destroy",
    "fir_cache_info", "fir_cache_forms",
    "fir_cache_tune", "fir_cache_forward", "fir_cache_backward",
    "fir_cache_sensitivity", "fir_cache_osem_project",
    "fir_histogram_events", "fir_rebin_weight",
    "fir_cache_job_workspace", "fir_osem_step_create", "fir_osem_step_run",
    "fir_osem_step_destroy",
)

JOB_MAX_IMAGES = 8


class VolDesc(C.Structure):
    """fir_vol_desc: the reference VolHeader (VolHeader.h:6-34)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double)]

    @classmethod
    def from_spec(cls, vol) -> "VolDesc":
        return cls(int(vol.size[0]), int(vol.size[1]), int(vol.size[2]),
                   float(vol.voxel[0]), float(vol.voxel[1]),
                   float(vol.voxel[2]), float(vol.offset[0]),
                   float(vol.offset[1]), float(vol.offset[2]))

    @property
    def n_voxels(self) -> int:
        return self.nx * self.ny * self.nz


class OsemJob(C.Structure):
    """fir_osem_job: up to 8 reconstructions on one path cache, one subset."""
    _fields_ = [("cache", C.c_void_p), ("subset", C.c_int),
                ("n_images", C.c_int),
                ("x", C.c_void_p * JOB_MAX_IMAGES),
                ("y", C.c_void_p * JOB_MAX_IMAGES),
                ("bias", C.c_void_p * JOB_MAX_IMAGES),
                ("bp", C.c_void_p * JOB_MAX_IMAGES),
                ("sens", C.c_void_p), ("workspace", C.c_void_p)]


class FirError(RuntimeError):
    pass


_lib = None


def load() -> C.CDLL:
    """Load libfir_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FirError(
            f"{LIB_PATH} is missing: build it with `python -m fir_amd.build`")
    lib = C.CDLL(LIB_PATH)
    lib.fir_last_error.restype = C.c_char_p
    lib.fir_proj_n_bins.restype = C.c_int64
    lib.fir_proj_n_bins.argtypes = [C.c_void_p]
    vp, i, i64, f = C.c_void_p, C.c_int, C.c_int64, C.c_float
    P = C.POINTER(VolDesc)
    sig = {
        "fir_ctx_create": [C.POINTER(C.c_void_p), i],
        "fir_ctx_destroy": [vp],
        "fir_set_stream": [vp, vp],
        "fir_synchronize": [vp],
        "fir_set_scanner": [vp, vp, i, vp, i],
        "fir_set_proj": [vp, i, i, i, i, i],
        "fir_set_walk_mode": [vp, i],
        "fir_set_uneven_subsets": [vp, i],
        "fir_set_lor_order": [vp, i],
        "fir_set_scatter_copies": [vp, i],
        "fir_fill": [vp, vp, i64, f],
        "fir_forward": [vp, P, vp, vp],
        "fir_forward_subset": [vp, P, vp, vp, i, i],
        "fir_backward": [vp, P, vp, vp, i],
        "fir_backward_subset": [vp, P, vp, vp, i, i],
        "fir_sensitivity": [vp, P, vp, i],
        "fir_osem_project": [vp, P, vp, vp, vp, vp, i, i],
        "fir_osem_update": [vp, P, vp, vp, vp, f, i, i],
        "fir_osem_subiter": [vp, P, vp, vp, vp, vp, vp, i, i, i, vp, f, i],
        "fir_convolve": [vp, P, vp, i, vp, f],
        "fir_convolve_noalloc": [vp, P, vp, i, vp, f],
        "fir_cut_circle": [vp, P, vp, i, f],
        "fir_mul_clamped": [vp, vp, vp, i64],
        "fir_div_clamped": [vp, vp, vp, i64],
        "fir_exp_clamped": [vp, vp, i64],
        "fir_hu_to_mu": [vp, vp, i64],
        "fir_path_dump": [vp, P, i64, i64, i, vp, vp, vp],
        "fir_path_count": [vp, P, i, i, C.POINTER(C.c_int64)],
        "fir_histogram_events": [vp, vp, i64, vp, vp, C.POINTER(C.c_int64)],
        "fir_rebin_weight": [vp, vp],
        "fir_poisson_nll_grad": [vp, P, vp, vp, vp, vp, vp],
        "fir_osem_project_batch": [vp, P, i, vp, vp, vp, vp, i, i],
        "fir_cache_create": [vp, P, i, i, C.POINTER(C.c_void_p)],
        "fir_cache_forms": [vp, C.POINTER(C.c_int)],
        "fir_cache_create_partial": [vp, P, i, i, i, i64,
                                     C.POINTER(C.c_void_p)],
        "fir_cache_subsets": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "fir_cache_destroy": [vp],
        "fir_cache_info": [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                           C.POINTER(C.c_int64)],
        "fir_cache_tune": [vp, i, i, i],
        "fir_set_cache_lanes": [vp, i, i],
        "fir_set_cache_zslices": [vp, i],
        "fir_set_cache_chunk": [vp, i64],
        "fir_set_cache_sheets": [vp, i],
        "fir_set_walk_sheets": [vp, i],
        "fir_cache_forward": [vp, vp, vp],
        "fir_cache_backward": [vp, vp, vp],
        "fir_cache_sensitivity": [vp, vp],
        "fir_cache_osem_project": [vp, i, vp, vp, vp, vp, i],
        "fir_cache_job_workspace": [vp, C.POINTER(C.c_int64)],
        "fir_osem_step_create": [i, vp, vp, f, C.POINTER(C.c_void_p)],
        "fir_osem_step_run": [vp],
        "fir_osem_step_destroy": [vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        msg = load().fir_last_error().decode(errors="replace")
        raise FirError(f"fir error {rc}: {msg}")


def _ptr(t) -> Optional[int]:
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise FirError("expected a device tensor")
    if not t.is_contiguous():
        raise FirError("expected a contiguous tensor")
    return t.data_ptr()


class Context:
    """One fir_ctx: scanner tables + sinogram geometry on one device.

    The operators mirror the reference's operator API (projections.h:9-24,
    reconAlgos.h:29-36, operations.h:7-22) on device tensors.
    """

    def __init__(self, device: int = 0, stream=None):
        self.lib = load()
        h = C.c_void_p()
        _check(self.lib.fir_ctx_create(C.byref(h), device))
        self.h = h
        self.device = device
        if stream is not None:
            self.set_stream(stream)

    def close(self):
        if getattr(self, "h", None):
            self.lib.fir_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up ------------------------------------------------------------
    def set_stream(self, stream):
        """Enqueue on a torch stream (its raw hipStream_t; 0 is the null
        stream)."""
        _check(self.lib.fir_set_stream(self.h, stream.cuda_stream or None))

    def synchronize(self):
        _check(self.lib.fir_synchronize(self.h))

    def set_scanner(self, crystal_xy: np.ndarray, slice_z: np.ndarray):
        xy = np.ascontiguousarray(crystal_xy, np.float64)
        z = np.ascontiguousarray(slice_z, np.float64)
        self._xy, self._z = xy, z
        _check(self.lib.fir_set_scanner(
            self.h, xy.ctypes.data, xy.size // 2, z.ctypes.data, z.size))

    def set_proj(self, n_rings, n_cpr, span, n_segments, n_tang):
        _check(self.lib.fir_set_proj(self.h, n_rings, n_cpr, span,
                                     n_segments, n_tang))

    def set_proj_spec(self, proj):
        self.set_proj(proj.n_rings, proj.n_crystals_per_ring,
                      proj.segment_span, proj.n_segments, proj.n_tang_coords)

    def set_walk_mode(self, mode: int):
        _check(self.lib.fir_set_walk_mode(self.h, mode))

    def set_lor_order(self, order: int):
        _check(self.lib.fir_set_lor_order(self.h, order))
        self._lor_order = order

    def set_scatter_copies(self, copies: int):
        _check(self.lib.fir_set_scatter_copies(self.h, copies))

    def set_cache_lanes(self, fwd_lanes: int, bp_lanes: int):
        """Lanes per LOR / per voxel the next path caches are laid out for."""
        _check(self.lib.fir_set_cache_lanes(self.h, fwd_lanes, bp_lanes))
        self._cache_lanes = (fwd_lanes, bp_lanes)

    def set_cache_zslices(self, zq: int):
        """z-slices interleaved in the gather layouts of the next caches."""
        _check(self.lib.fir_set_cache_zslices(self.h, zq))
        self._cache_zq = zq

    def set_walk_sheets(self, enable: bool):
        """Whether osem_project runs family by family with LDS sheets."""
        _check(self.lib.fir_set_walk_sheets(self.h, int(enable)))

    def set_cache_sheets(self, enable: bool):
        """Whether the next path matrices also get the sheet form."""
        _check(self.lib.fir_set_cache_sheets(self.h, int(enable)))

    def set_cache_chunk(self, max_elements: int):
        """Largest number of path elements one transpose step of the next
        cache builds sorts (24 bytes of temporaries per element)."""
        _check(self.lib.fir_set_cache_chunk(self.h, int(max_elements)))

    def set_uneven_subsets(self, allow: bool):
        _check(self.lib.fir_set_uneven_subsets(self.h, int(allow)))

    @property
    def n_bins(self) -> int:
        return int(self.lib.fir_proj_n_bins(self.h))

    # -- operators (device tensors) -----------------------------------------
    def forward(self, vol: VolDesc, img, sino, subset=None, n_subsets=0):
        if subset is None:
            _check(self.lib.fir_forward(self.h, C.byref(vol), _ptr(img),
                                        _ptr(sino)))
        else:
            _check(self.lib.fir_forward_subset(self.h, C.byref(vol), _ptr(img),
                                               _ptr(sino), subset, n_subsets))

    def backward(self, vol: VolDesc, sino, img, n_subsets=1):
        _check(self.lib.fir_backward(self.h, C.byref(vol), _ptr(sino),
                                     _ptr(img), n_subsets))

    def backward_subset(self, vol: VolDesc, sino, img, subset, n_subsets):
        _check(self.lib.fir_backward_subset(self.h, C.byref(vol), _ptr(sino),
                                            _ptr(img), subset, n_subsets))

    def sensitivity(self, vol: VolDesc, sens, n_subsets=1):
        _check(self.lib.fir_sensitivity(self.h, C.byref(vol), _ptr(sens),
                                        n_subsets))

    def osem_project(self, vol, x, y, bias, bp, subset, n_subsets):
        _check(self.lib.fir_osem_project(self.h, C.byref(vol), _ptr(x),
                                         _ptr(y), _ptr(bias), _ptr(bp),
                                         subset, n_subsets))

    def osem_project_batch(self, vol, xs, ys, biases, bps, subset,
                           n_subsets):
        """One projector pass for several reconstructions of one geometry."""
        n = len(xs)
        arr = C.c_void_p * n
        px = arr(*[_ptr(t) for t in xs])
        py = arr(*[_ptr(t) for t in ys])
        pb = None if biases is None else arr(*[_ptr(t) for t in biases])
        pp = arr(*[_ptr(t) for t in bps])
        _check(self.lib.fir_osem_project_batch(self.h, C.byref(vol), n, px,
                                               py, pb, pp, subset, n_subsets))

    def osem_update(self, vol, x, bp, sens_s, cut=0.0, apply_cut=True,
                    reset=True):
        _check(self.lib.fir_osem_update(self.h, C.byref(vol), _ptr(x),
                                        _ptr(bp), _ptr(sens_s), cut,
                                        int(apply_cut), int(reset)))

    def osem_subiter(self, vol, x, y, bias, sens_s, bp, subset, n_subsets,
                     convolve=False, fwhm=(0.0, 0.0, 0.0), cut=0.0,
                     reset=True):
        fw = (C.c_float * 3)(*fwhm)
        _check(self.lib.fir_osem_subiter(
            self.h, C.byref(vol), _ptr(x), _ptr(y), _ptr(bias), _ptr(sens_s),
            _ptr(bp), subset, n_subsets, int(convolve), fw, cut, int(reset)))

    def convolve(self, vol, img, fwhm, cut=0.0, n_frames=1):
        fw = (C.c_float * 3)(*fwhm)
        _check(self.lib.fir_convolve(self.h, C.byref(vol), _ptr(img),
                                     n_frames, fw, cut))

    def convolve_noalloc(self, vol, img, fwhm, cut=0.0, n_frames=1):
        """fir_convolve through the library's static scratch: allocates no
        device memory (frames of up to 2^23 voxels)."""
        fw = (C.c_float * 3)(*fwhm)
        _check(self.lib.fir_convolve_noalloc(self.h, C.byref(vol), _ptr(img),
                                             n_frames, fw, cut))

    def cut_circle(self, vol, img, cut, n_frames=1):
        _check(self.lib.fir_cut_circle(self.h, C.byref(vol), _ptr(img),
                                       n_frames, cut))

    def fill(self, t, value: float):
        _check(self.lib.fir_fill(self.h, _ptr(t), t.numel(), value))

    def mul_clamped(self, a, b):
        _check(self.lib.fir_mul_clamped(self.h, _ptr(a), _ptr(b), a.numel()))

    def div_clamped(self, a, b):
        _check(self.lib.fir_div_clamped(self.h, _ptr(a), _ptr(b), a.numel()))

    def exp_clamped(self, a):
        _check(self.lib.fir_exp_clamped(self.h, _ptr(a), a.numel()))

    def hu_to_mu(self, a):
        _check(self.lib.fir_hu_to_mu(self.h, _ptr(a), a.numel()))

    def poisson_nll_grad(self, vol, x, y, r, grad, loss):
        """loss[0] = sum(l - y log l), grad = A^T(1 - y/l), l = A x + r."""
        _check(self.lib.fir_poisson_nll_grad(self.h, C.byref(vol), _ptr(x),
                                             _ptr(y), _ptr(r), _ptr(grad),
                                             _ptr(loss)))

    def path_count(self, vol, subset=0, n_subsets=0) -> int:
        """Total path elements over one subset (n_subsets=0: all bins)."""
        out = C.c_int64()
        _check(self.lib.fir_path_count(self.h, C.byref(vol), subset,
                                       n_subsets, C.byref(out)))
        return int(out.value)

    def histogram_events(self, events, sino, weights=None,
                         count: bool = False) -> Optional[int]:
        """sino += histogram of list-mode events (int32 [n, 4]: ring1,
        crystal1, ring2, crystal2) (ProjData::getBinCoordinates +
        incrementBin); with count=True returns the number of events kept."""
        if events.dtype.itemsize != 4 or events.dim() != 2 or \
                events.shape[1] != 4:
            raise FirError("events must be int32 [n, 4]")
        out = C.c_int64() if count else None
        _check(self.lib.fir_histogram_events(
            self.h, _ptr(events), events.shape[0], _ptr(weights), _ptr(sino),
            C.byref(out) if count else None))
        return int(out.value) if count else None

    def rebin_weight(self, sino):
        """Span > 1 fused-bin weights, in place (ProjData::rebinWeight)."""
        _check(self.lib.fir_rebin_weight(self.h, _ptr(sino)))

    def path_dump(self, vol, bin0, n_bins, max_per_lor, counts, idx, ln):
        _check(self.lib.fir_path_dump(self.h, C.byref(vol), bin0, n_bins,
                                      max_per_lor, _ptr(counts), _ptr(idx),
                                      _ptr(ln)))

    def path_cache(self, vol: VolDesc, n_subsets: int = 1,
                   len_format: int = FIR_LEN_SNORM16,
                   row_order: Optional[int] = None,
                   max_subsets: Optional[int] = None,
                   max_bytes: Optional[int] = None) -> "PathCache":
        """Build a path cache. row_order (FIR_ORDER_* flags) sets the order
        of the cached rows within a view (the context's LOR order at
        creation); by default the context's current order is used. With
        max_subsets or max_bytes the cache is partial
        (fir_cache_create_partial): it holds subsets 0 .. n_cached-1 only,
        as many as fit (0 or None: no cap, no limit beyond free memory)."""
        partial = None
        if max_subsets is not None or max_bytes is not None:
            partial = (int(max_subsets or 0), int(max_bytes or 0))
        if row_order is None:
            return PathCache(self, vol, n_subsets, len_format, partial)
        old = getattr(self, "_lor_order", 0)
        self.set_lor_order(row_order)
        try:
            return PathCache(self, vol, n_subsets, len_format, partial)
        finally:
            self.set_lor_order(old)


def _ptr_array(ts):
    return (C.c_void_p * len(ts))(*[_ptr(t) for t in ts])


class PathCache:
    """fir_cache: the Siddon paths of one volume kept in device memory (the
    LOR-major matrix and its voxel-major transpose for one subset split),
    an extension of the reference's LORCache (LORCache.h:5-47)."""

    def __init__(self, ctx: Context, vol: VolDesc, n_subsets: int = 1,
                 len_format: int = FIR_LEN_SNORM16, partial=None):
        self.ctx, self.lib, self.vol = ctx, ctx.lib, vol
        self.n_subsets, self.len_format = n_subsets, len_format
        h = C.c_void_p()
        if partial is None:
            _check(self.lib.fir_cache_create(ctx.h, C.byref(vol), n_subsets,
                                             len_format, C.byref(h)))
        else:
            _check(self.lib.fir_cache_create_partial(
                ctx.h, C.byref(vol), n_subsets, len_format, partial[0],
                partial[1], C.byref(h)))
        self.h = h
        s, k = C.c_int(), C.c_int()
        _check(self.lib.fir_cache_subsets(h, C.byref(s), C.byref(k)))
        self.n_cached = int(k.value)

    def close(self):
        if getattr(self, "h", None):
            self.lib.fir_cache_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        r, e, b = C.c_int64(), C.c_int64(), C.c_int64()
        _check(self.lib.fir_cache_info(self.h, C.byref(r), C.byref(e),
                                       C.byref(b)))
        return dict(rows=int(r.value), elements=int(e.value),
                    device_bytes=int(b.value))

    def forms(self) -> dict:
        """Which forms of the path matrix the cache holds
        (fir_cache_forms)."""
        f = C.c_int()
        _check(self.lib.fir_cache_forms(self.h, C.byref(f)))
        return dict(sheets=bool(f.value & FIR_FORM_SHEETS),
                    family_lists=bool(f.value & FIR_FORM_FAMILY_LISTS),
                    slabs=bool(f.value & FIR_FORM_SLABS))

    def tune(self, fwd_lanes: int = 16, bp_lanes: int = 16, flags: int = 0):
        _check(self.lib.fir_cache_tune(self.h, fwd_lanes, bp_lanes, flags))

    def forward(self, img, sino):
        _check(self.lib.fir_cache_forward(self.h, _ptr(img), _ptr(sino)))

    def backward(self, sino, frames):
        _check(self.lib.fir_cache_backward(self.h, _ptr(sino), _ptr(frames)))

    def sensitivity(self, frames):
        _check(self.lib.fir_cache_sensitivity(self.h, _ptr(frames)))

    def osem_project(self, xs, ys, biases, bps, subset):
        n = len(xs)
        pb = None if biases is None else _ptr_array(biases)
        _check(self.lib.fir_cache_osem_project(
            self.h, n, _ptr_array(xs), _ptr_array(ys), pb, _ptr_array(bps),
            subset))

    def job_workspace(self) -> int:
        """Device bytes of one fir_osem_job workspace on this cache."""
        out = C.c_int64()
        _check(self.lib.fir_cache_job_workspace(self.h, C.byref(out)))
        return int(out.value)


class OsemStep:
    """fir_osem_step: one sub-iteration of many jobs (dicts with keys
    cache, subset, x, y, bp, sens, workspace and optionally bias; x, y, bp
    and bias lists of up to 8 device tensors), every phase of all jobs in
    one launch. The tensors must stay alive and in place while the step
    exists (its tables hold their addresses)."""

    def __init__(self, jobs, fwhm=(0.0, 0.0, 0.0), cut: float = 0.0):
        self.lib = load()
        arr = (OsemJob * len(jobs))()
        self._keep = []
        for k, jd in enumerate(jobs):
            j = arr[k]
            j.cache = jd["cache"].h
            j.subset = int(jd["subset"])
            n = len(jd["x"])
            if not 1 <= n <= JOB_MAX_IMAGES:
                raise ValueError(f"a job holds 1 to {JOB_MAX_IMAGES} images")
            j.n_images = n
            bias = jd.get("bias") or [None] * n
            for b in range(n):
                j.x[b] = _ptr(jd["x"][b])
                j.y[b] = _ptr(jd["y"][b])
                j.bp[b] = _ptr(jd["bp"][b])
                j.bias[b] = _ptr(bias[b])
            j.sens = _ptr(jd["sens"])
            j.workspace = _ptr(jd["workspace"])
            self._keep.append(jd)
        fw = (C.c_float * 3)(*fwhm)
        h = C.c_void_p()
        _check(self.lib.fir_osem_step_create(len(jobs), arr, fw, float(cut),
                                             C.byref(h)))
        self.h = h

    def run(self):
        _check(self.lib.fir_osem_step_run(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.fir_osem_step_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_libm = None


def _sincos(angle: float):
    """(sin, cos) from the C library's sincos: the reference's
    rotateXY (ScannerData.cc:234-246) calls cos and sin on the same angle,
    which an optimising C++ compiler merges into one sincos call; separate
    cos/sin can differ from it in the last place for some angles."""
    global _libm
    if _libm is None:
        _libm = C.CDLL("libm.so.6")
        _libm.sincos.argtypes = [C.c_double, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
        _libm.sincos.restype = None
    s, c = C.c_double(), C.c_double()
    _libm.sincos(angle, C.byref(s), C.byref(c))
    return s.value, c.value


def scanner_tables(scanner) -> tuple:
    """Crystal (x, y) per angular index and slice z, computed as the
    reference ScannerData does (ScannerHeader.cc:305-342,
    ScannerData.cc:234-377). Host-side geometry, run once per scanner."""
    cd = list(scanner.crystal_dimensions)
    crep = list(scanner.crystal_repeat_numbers)
    cgap = list(scanner.inter_crystal_distance)
    md = list(scanner.module_dimensions)
    mrep = list(scanner.module_repeat_numbers)
    mgap = list(scanner.inter_module_distance)
    rd = list(scanner.rsector_dimensions)
    rrep = int(scanner.rsector_repeat_number)
    rin = float(scanner.rsector_inner_radius)
    if md[0] == 0.0:
        md[0] = cd[0]
    if md[1] == 0.0:
        md[1] = cd[1] * crep[0] + cgap[0] * (crep[0] - 1)
    if md[2] == 0.0:
        md[2] = cd[2] * crep[1] + cgap[1] * (crep[1] - 1)
    if rd[0] == 0:
        rd[0] = md[0]
    if rd[1] == 0:
        rd[1] = md[1] * mrep[0] + mgap[0] * (mrep[0] - 1)
    if rd[2] == 0:
        rd[2] = md[2] * mrep[1] + mgap[1] * (mrep[1] - 1)
    cvy, cvz = cd[1] + cgap[0], cd[2] + cgap[1]
    mvy, mvz = md[1] + mgap[0], md[2] + mgap[1]
    tx = rin + rd[0] / 2
    n_cpr = crep[0] * mrep[0] * rrep
    n_rings = crep[1] * mrep[1]
    coff = mrep[0] * crep[0] // 2
    per = mrep[0] * crep[0]
    y0 = -((mrep[0] - 1) * mvy + (crep[0] - 1) * cvy) / 2.0
    ys = [m * mvy + c * cvy + y0 for m in range(mrep[0]) for c in range(crep[0])]
    dang = 2 * 3.141592653589793 / rrep
    first = n_cpr - coff
    xy = np.zeros(2 * n_cpr)
    for i in range(first):
        k = i + coff
        yy = ys[k % per]
        ang = (k // per) * dang
        sa, ca = _sincos(ang)
        xy[2 * i] = tx * ca - yy * sa
        xy[2 * i + 1] = tx * sa + yy * ca
    for i in range(first, n_cpr):
        xy[2 * i] = tx
        xy[2 * i + 1] = ys[i - first]
    perz = mrep[1] * crep[1]
    z0 = -((mrep[1] - 1) * mvz + (crep[1] - 1) * cvz) / 2.0
    rz = [m * mvz + c * cvz + z0 for m in range(mrep[1]) for c in range(crep[1])]
    ns = 2 * n_rings - 1
    z = np.array([rz[s // 2] if s % 2 == 0 else (rz[s // 2] + rz[(s + 1) // 2]) / 2.0
                  for s in range(ns)])
    return xy, z


def context_for(workload, device: int = 0, geometric: bool = False) -> Context:
    """Context with the scanner and sinogram geometry of a synth.Workload."""
    ctx = Context(device)
    try:
        import torch
        ctx.set_stream(torch.cuda.current_stream(device))
    except Exception:  # no torch device: keep the context's own stream
        pass
    xy, z = scanner_tables(workload.scanner)
    ctx.set_scanner(xy, z)
    ctx.set_proj_spec(workload.proj)
    if geometric:
        ctx.set_walk_mode(FIR_WALK_GEOMETRIC)
    return ctx
