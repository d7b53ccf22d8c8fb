"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product path (fir_amd)
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MAX_SEG = 1024


class OrcVol(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double)]


class OrcProj(C.Structure):
    _fields_ = [("n_rings", C.c_int), ("n_cpr", C.c_int), ("span", C.c_int),
                ("n_seg", C.c_int), ("n_tang", C.c_int),
                ("seg_off", C.c_int), ("tang_off", C.c_int),
                ("n_views", C.c_int), ("half_span", C.c_int),
                ("max_ring_diff", C.c_int), ("n_bins", C.c_longlong),
                ("n_axial", C.c_int * MAX_SEG),
                ("seg_base", C.c_longlong * MAX_SEG)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (make -C oracle)")
        _lib = C.CDLL(path)
        assert _lib.orc_sizeof_proj() == C.sizeof(OrcProj)
        assert _lib.orc_sizeof_vol() == C.sizeof(OrcVol)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def vol_desc(vol) -> OrcVol:
    """OrcVol from a fir_amd.interfile.VolSpec."""
    return OrcVol(int(vol.size[0]), int(vol.size[1]), int(vol.size[2]),
                  float(vol.voxel[0]), float(vol.voxel[1]),
                  float(vol.voxel[2]), float(vol.offset[0]),
                  float(vol.offset[1]), float(vol.offset[2]))


def proj_desc(proj) -> OrcProj:
    p = OrcProj()
    rc = lib().orc_proj_fill(C.byref(p), proj.n_rings,
                             proj.n_crystals_per_ring, proj.segment_span,
                             proj.n_segments, proj.n_tang_coords)
    if rc:
        raise ValueError(f"invalid projection header (code {rc})")
    return p


def proj_fill(n_rings, n_cpr, span, n_seg, n_tang):
    p = OrcProj()
    rc = lib().orc_proj_fill(C.byref(p), n_rings, n_cpr, span, n_seg, n_tang)
    return rc, p


def scanner_tables(scanner):
    hdr = scanner.header_vector()
    n_cpr = scanner.n_crystals_per_ring
    n_rings = scanner.n_rings
    xy = np.zeros(2 * n_cpr, np.float64)
    z = np.zeros(2 * n_rings - 1, np.float64)
    lib().orc_scanner_tables(_p(hdr), _p(xy), _p(z), None, None)
    return xy, z


def crystal_axial(p: OrcProj, seg, axial):
    a, b = C.c_int(), C.c_int()
    lib().orc_crystal_axial(C.byref(p), seg, axial, C.byref(a), C.byref(b))
    return a.value, b.value


def crystal_ang(p: OrcProj, view, tang):
    a, b = C.c_int(), C.c_int()
    lib().orc_crystal_ang(C.byref(p), view, tang, C.byref(a), C.byref(b))
    return a.value, b.value


def bin_coords(p: OrcProj, r1, c1, r2, c2):
    """(seg, view, axial, tang) of a crystal pair, or None if outside."""
    sg, v, a, t = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    ok = lib().orc_bin_coords(C.byref(p), int(r1), int(c1), int(r2), int(c2),
                              C.byref(sg),
                              C.byref(v), C.byref(a), C.byref(t))
    return (sg.value, v.value, a.value, t.value) if ok else None


def bin_index(p: OrcProj, seg, view, axial, tang):
    lib().orc_bin_index.restype = C.c_longlong
    return lib().orc_bin_index(C.byref(p), int(seg), int(view), int(axial),
                               int(tang))


def histogram_events(proj, events, weights=None, sino=None):
    """Sinogram counts of list-mode events (int32 [n, 4]: r1, c1, r2, c2);
    returns (sino float32, number counted)."""
    p = proj if isinstance(proj, OrcProj) else proj_desc(proj)
    ev = np.ascontiguousarray(events, np.int32)
    if sino is None:
        sino = np.zeros(p.n_bins, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    lib().orc_histogram_events.restype = C.c_longlong
    n = lib().orc_histogram_events(C.byref(p), _p(ev),
                                   C.c_longlong(ev.shape[0]),
                                   None if w is None else _p(w), _p(sino))
    return sino, int(n)


def rebin_weight(proj, sino):
    """In place: span > 1 fused-bin weights (ProjData::rebinWeight)."""
    p = proj if isinstance(proj, OrcProj) else proj_desc(proj)
    assert sino.dtype == np.float32 and sino.flags.c_contiguous
    lib().orc_rebin_weight(C.byref(p), _p(sino))
    return sino


def path(vol, p1, p2, geometric=False):
    """Siddon path of one segment: (idx int32[n], len float64[n]) or None
    for a miss."""
    v = vol_desc(vol) if not isinstance(vol, OrcVol) else vol
    cap = v.nx + v.ny + v.nz + 16
    idx = np.zeros(cap, np.int32)
    ln = np.zeros(cap, np.float64)
    a = np.asarray(p1, np.float64)
    b = np.asarray(p2, np.float64)
    n = lib().orc_path(C.byref(v), int(geometric), _p(a), _p(b), _p(idx),
                       _p(ln), cap)
    if n < 0:
        return None
    return idx[:n].copy(), ln[:n].copy()


def path_dump(vol, proj, xy, z, b0=0, nb=None, geometric=False, cap=None):
    """counts[nb] (-1 = miss), idx[total], len[total] for bins [b0, b0+nb)."""
    v = vol_desc(vol)
    p = proj_desc(proj)
    if nb is None:
        nb = p.n_bins - b0
    counts = np.zeros(nb, np.int32)
    if cap is None:
        cap = nb * (v.nx + v.ny + v.nz)
    idx = np.zeros(cap, np.int32)
    ln = np.zeros(cap, np.float64)
    lib().orc_path_dump.restype = C.c_longlong
    tot = lib().orc_path_dump(C.byref(v), int(geometric), C.byref(p),
                              _p(xy), _p(z), C.c_longlong(b0),
                              C.c_longlong(nb), _p(counts), _p(idx), _p(ln),
                              C.c_longlong(cap))
    if tot > cap:
        raise RuntimeError("path dump capacity exceeded")
    return counts, idx[:tot].copy(), ln[:tot].copy()


def _sfx(dtype):
    return "_f64" if np.dtype(dtype) == np.float64 else ""


def forward(vol, proj, xy, z, img, geometric=False, nthreads=0):
    v, p = vol_desc(vol), proj_desc(proj)
    img = np.ascontiguousarray(img).ravel()
    sino = np.zeros(p.n_bins, img.dtype)
    getattr(lib(), "orc_forward" + _sfx(img.dtype))(
        C.byref(v), int(geometric), C.byref(p), _p(xy), _p(z), _p(img),
        _p(sino), nthreads)
    return sino


def forward_bins(vol, proj, xy, z, img, bins, geometric=False, nthreads=0):
    """Line integrals of the listed bins (all-bins order), as forward()
    computes them; dtype of img (float32 or float64)."""
    v, p = vol_desc(vol), proj_desc(proj)
    img = np.ascontiguousarray(img).ravel()
    b = np.ascontiguousarray(bins, np.int64)
    out = np.zeros(b.size, img.dtype)
    getattr(lib(), "orc_forward_bins" + _sfx(img.dtype))(
        C.byref(v), int(geometric), C.byref(p), _p(xy), _p(z), _p(img),
        _p(b), C.c_longlong(b.size), _p(out), nthreads)
    return out


def backward_bins(vol, proj, xy, z, bins, vals, geometric=False,
                  nthreads=0):
    """Backprojection of the listed bins only (values vals), one frame."""
    v, p = vol_desc(vol), proj_desc(proj)
    vals = np.ascontiguousarray(vals).ravel()
    b = np.ascontiguousarray(bins, np.int64)
    img = np.zeros(v.nx * v.ny * v.nz, vals.dtype)
    getattr(lib(), "orc_backward_bins" + _sfx(vals.dtype))(
        C.byref(v), int(geometric), C.byref(p), _p(xy), _p(z), _p(b),
        _p(vals), C.c_longlong(b.size), _p(img), nthreads)
    return img


def backward(vol, proj, xy, z, sino, n_subsets=1, geometric=False,
             nthreads=0):
    v, p = vol_desc(vol), proj_desc(proj)
    sino = np.ascontiguousarray(sino).ravel()
    nvox = v.nx * v.ny * v.nz
    img = np.zeros(nvox * n_subsets, sino.dtype)
    getattr(lib(), "orc_backward" + _sfx(sino.dtype))(
        C.byref(v), int(geometric), C.byref(p), _p(xy), _p(z), _p(sino),
        _p(img), n_subsets, nthreads)
    return img


def sensitivity(vol, proj, xy, z, n_subsets=1, dtype=np.float32,
                geometric=False, nthreads=0):
    v, p = vol_desc(vol), proj_desc(proj)
    nvox = v.nx * v.ny * v.nz
    img = np.zeros(nvox * n_subsets, dtype)
    getattr(lib(), "orc_sensitivity" + _sfx(dtype))(
        C.byref(v), int(geometric), C.byref(p), _p(xy), _p(z), _p(img),
        n_subsets, nthreads)
    return img


def osem(vol, proj, xy, z, y, sens, x0, n_iter, n_subsets, cut=0.0,
         conv_interval=0, fwhm=(0.0, 0.0, 0.0), bias=None, geometric=False,
         nthreads=0):
    v, p = vol_desc(vol), proj_desc(proj)
    dt = np.asarray(y).dtype
    y = np.ascontiguousarray(y, dt).ravel()
    sens = np.ascontiguousarray(sens, dt).ravel()
    x = np.ascontiguousarray(x0, dt).ravel().copy()
    fw = np.asarray(fwhm, np.float32)
    b = None if bias is None else np.ascontiguousarray(bias, dt).ravel()
    fn = getattr(lib(), "orc_osem" + _sfx(dt))
    rc = fn(C.byref(v), int(geometric), C.byref(p), _p(xy), _p(z), _p(y),
            None if b is None else _p(b), _p(sens), _p(x), n_iter, n_subsets,
            C.c_float(cut), conv_interval, _p(fw), nthreads)
    if rc:
        raise ValueError("number of subsets must divide the number of views")
    return x


def osem_reso(vol, proj, xy, z, y, sens, x0, n_iter, n_subsets, cut=0.0,
              conv_interval=0, fwhm=(0.0, 0.0, 0.0), bias=None,
              geometric=False, nthreads=0):
    """reconAlgos::OSEM_ResoReco (reconAlgos.cc:203-367); sens is not
    modified (a copy is convolved)."""
    v, p = vol_desc(vol), proj_desc(proj)
    dt = np.asarray(y).dtype
    y = np.ascontiguousarray(y, dt).ravel()
    sens = np.ascontiguousarray(sens, dt).ravel().copy()
    x = np.ascontiguousarray(x0, dt).ravel().copy()
    fw = np.asarray(fwhm, np.float32)
    b = None if bias is None else np.ascontiguousarray(bias, dt).ravel()
    fn = getattr(lib(), "orc_osem_reso" + _sfx(dt))
    rc = fn(C.byref(v), int(geometric), C.byref(p), _p(xy), _p(z), _p(y),
            None if b is None else _p(b), _p(sens), _p(x), n_iter, n_subsets,
            C.c_float(cut), conv_interval, _p(fw), nthreads)
    if rc:
        raise ValueError("number of subsets must divide the number of views")
    return x


def convolve(vol, img, fwhm, cut=0.0):
    v = vol_desc(vol)
    x = np.ascontiguousarray(img).ravel().copy()
    fw = np.asarray(fwhm, np.float32)
    getattr(lib(), "orc_convolve" + _sfx(x.dtype))(
        C.byref(v), _p(x), _p(fw), C.c_float(cut))
    return x


def gauss_taps(vol, fwhm):
    v = vol_desc(vol)
    kx = np.zeros(64, np.float32)
    ky = np.zeros(64, np.float32)
    kz = np.zeros(64, np.float32)
    h = np.zeros(3, np.int32)
    fw = np.asarray(fwhm, np.float32)
    rc = lib().orc_gauss_taps(C.byref(v), _p(fw), _p(kx), _p(ky), _p(kz),
                              _p(h))
    if rc:
        raise ValueError("kernel too wide")
    return (kx[:2 * h[0] + 1].copy(), ky[:2 * h[1] + 1].copy(),
            kz[:2 * h[2] + 1].copy())


def cut_circle(vol, img, cut):
    v = vol_desc(vol)
    x = np.ascontiguousarray(img).ravel().copy()
    getattr(lib(), "orc_cut_circle" + _sfx(x.dtype))(C.byref(v), _p(x),
                                                     C.c_float(cut))
    return x


def mul_clamped(a, b):
    a = np.ascontiguousarray(a).ravel().copy()
    b = np.ascontiguousarray(b, a.dtype).ravel()
    getattr(lib(), "orc_mul_clamped" + _sfx(a.dtype))(
        _p(a), _p(b), C.c_longlong(a.size))
    return a


def div_clamped(a, b):
    a = np.ascontiguousarray(a).ravel().copy()
    b = np.ascontiguousarray(b, a.dtype).ravel()
    getattr(lib(), "orc_div_clamped" + _sfx(a.dtype))(
        _p(a), _p(b), C.c_longlong(a.size))
    return a


def exp_clamped(a):
    a = np.ascontiguousarray(a).ravel().copy()
    getattr(lib(), "orc_exp_clamped" + _sfx(a.dtype))(_p(a),
                                                      C.c_longlong(a.size))
    return a


def hu_to_mu(a):
    a = np.ascontiguousarray(a, np.float32).ravel().copy()
    lib().orc_hu_to_mu(_p(a), C.c_longlong(a.size))
    return a
