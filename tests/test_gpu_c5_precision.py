"""The 64-ring scanner (64 rings x 672 crystals, 127 segments, 225
tangential bins: 309 657 600 LORs; 256x256x128 image at 1.6 mm) against the
all-FP64 oracle. The reference itself is invalid here (its int-truncated
step extents walk 1.6 mm voxels as 1 mm, Siddon.cc:35-38, and its 16-bit
crystal ids overflow, LORCache.h:5-10), so the walk is geometric and the
checker is the oracle in FP64 throughout.

* forward projection of the seeded phantom: every 997th bin (310 590 bins)
  against the oracle's FP64 line integrals of the same bins, computed here;
* backprojection of a sparse sinogram (those bins only, hash counts)
  against the oracle's FP64 backprojection of the same bins, computed here;
* backprojection of the full hash-count sinogram, the 24-subset
  sensitivity and OSEM 2 x 24 from a uniform image (the plain path: FP64
  walk, FP32 values) against the FP64 oracle's results, kept as block sums
  and central planes in tests/golden/c5_fp64_golden.npz (made by
  tests/golden/make_golden.py c5fp64);
* a partial path cache (2 of 24 subsets, snorm16 lengths with a per-LOR
  FP32 scale): its forward projection and backprojection of the cached
  subsets' sampled bins against the oracle.
"""
import os

import numpy as np
import pytest

import pyoracle as O
from conftest import GOLDEN, rel_l2, rel_max
from fir_amd import lib as firlib
from fir_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SAMPLE = 997


@pytest.fixture(scope="module")
def c5():
    w = synth.c5()
    ctx = firlib.context_for(w, geometric=True)
    v = firlib.VolDesc.from_spec(w.vol)
    xy, z = O.scanner_tables(w.scanner)
    act, _ = synth.phantom(w)
    nb = ctx.n_bins
    bins = np.arange(0, nb, SAMPLE, dtype=np.int64)
    y = synth.hash_counts(nb)
    d = dict(w=w, ctx=ctx, v=v, xy=xy, z=z, act=act.ravel(), bins=bins,
             y=torch.from_numpy(y).cuda(), y_bins=y[bins].astype(np.float64),
             shape=tuple(reversed(w.vol.size)))
    yield d
    ctx.close()


def _views_of(proj, bins):
    """View of each bin (segments in order, each [view][axial][tang],
    ProjData.h:16-21)."""
    p = O.proj_desc(proj)
    na = np.asarray(p.n_axial[:p.n_seg], np.int64)
    base = np.concatenate([[0], np.cumsum(na * p.n_views * p.n_tang)])
    k = np.searchsorted(base, bins, side="right") - 1
    return (bins - base[k]) // p.n_tang // na[k]


def _sparse(c5, bins, vals):
    s = torch.zeros(c5["ctx"].n_bins, device="cuda")
    s[torch.from_numpy(bins).cuda()] = torch.from_numpy(
        vals.astype(np.float32)).cuda()
    return s


def test_c5_forward_samples_vs_fp64(c5):
    w, ctx, v = c5["w"], c5["ctx"], c5["v"]
    sino = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v, torch.from_numpy(c5["act"]).cuda(), sino)
    got = sino[torch.from_numpy(c5["bins"]).cuda()].cpu().numpy()
    ref = O.forward_bins(w.vol, w.proj, c5["xy"], c5["z"],
                         c5["act"].astype(np.float64), c5["bins"],
                         geometric=True)
    big = ref > 1e-3 * ref.max()
    e2 = rel_l2(got, ref)
    em = float((np.abs(got - ref)[big] / ref[big]).max())
    print(f"C5 forward, {ref.size} bins: rel L2 {e2:.2e}, max over "
          f"significant bins {em:.2e}", flush=True)
    assert big.sum() > 50_000
    assert e2 < 1e-5 and em < 1e-5


def test_c5_sparse_backprojection_vs_fp64(c5):
    w, ctx, v = c5["w"], c5["ctx"], c5["v"]
    img = torch.zeros(v.n_voxels, device="cuda")
    ctx.backward(v, _sparse(c5, c5["bins"], c5["y_bins"]), img, 1)
    ref = O.backward_bins(w.vol, w.proj, c5["xy"], c5["z"], c5["bins"],
                          c5["y_bins"], geometric=True)
    got = img.cpu().numpy()
    print(f"C5 sparse backprojection: rel L2 {rel_l2(got, ref):.2e}, rel "
          f"max {rel_max(got, ref):.2e}", flush=True)
    assert rel_l2(got, ref) < 1e-5 and rel_max(got, ref) < 1e-5


def _summary_errors(img, golden, prefix, shape):
    s = synth.image_summary(img, shape)
    return {k: rel_l2(s[k], golden[prefix + k]) for k in s}


def test_c5_full_results_vs_fp64_golden(c5):
    g = dict(np.load(os.path.join(GOLDEN, "c5_fp64_golden.npz")))
    w, ctx, v = c5["w"], c5["ctx"], c5["v"]
    n_it, n_sub = (int(t) for t in g["schedule"])
    nv = v.n_voxels
    y = c5["y"]
    bp = torch.zeros(nv, device="cuda")
    ctx.backward(v, y, bp, 1)
    sens = torch.zeros(n_sub * nv, device="cuda")
    ctx.sensitivity(v, sens, n_sub)
    x = torch.ones(nv, device="cuda")
    acc = torch.zeros(nv, device="cuda")
    total = n_it * n_sub
    for it in range(n_it):
        for s in range(n_sub):
            sub = it * n_sub + s + 1
            ctx.osem_subiter(v, x, y, None, sens[s * nv:(s + 1) * nv], acc,
                             s, n_sub, reset=sub != total)
    torch.cuda.synchronize()
    shape = c5["shape"]
    res = {"bp": (_summary_errors(bp.cpu().numpy(), g, "bp_", shape), 1e-5),
           "sens": (_summary_errors(sens.view(n_sub, -1).sum(0).cpu().numpy(),
                                    g, "sens_", shape), 1e-5),
           "osem": (_summary_errors(x.cpu().numpy(), g, "osem_", shape),
                    1e-4)}
    for name, (errs, bar) in res.items():
        print(f"C5 {name} vs FP64: " + ", ".join(
            f"{k} {e:.2e}" for k, e in errs.items()), flush=True)
        assert max(errs.values()) < bar, (name, errs)


def test_c5_partial_cache_vs_fp64(c5):
    w, ctx, v = c5["w"], c5["ctx"], c5["v"]
    n_sub = w.n_subsets
    cache = ctx.path_cache(v, n_sub, firlib.FIR_LEN_SNORM16,
                           row_order=firlib.CACHE_ROW_ORDER, max_subsets=2)
    k = cache.n_cached
    assert 1 <= k <= 2
    bins = c5["bins"]
    keep = _views_of(w.proj, bins) % n_sub < k
    cb, cy = bins[keep], c5["y_bins"][keep]
    sino = torch.zeros(ctx.n_bins, device="cuda")
    cache.forward(torch.from_numpy(c5["act"]).cuda(), sino)
    got = sino[torch.from_numpy(cb).cuda()].cpu().numpy()
    ref = O.forward_bins(w.vol, w.proj, c5["xy"], c5["z"],
                         c5["act"].astype(np.float64), cb, geometric=True)
    big = ref > 1e-3 * ref.max()
    ef = rel_l2(got, ref)
    em = float((np.abs(got - ref)[big] / ref[big]).max())
    frames = torch.zeros(n_sub * v.n_voxels, device="cuda")
    cache.backward(_sparse(c5, cb, cy), frames)
    bp = frames.view(n_sub, -1)[:k].sum(0).cpu().numpy()
    cache.close()
    rbp = O.backward_bins(w.vol, w.proj, c5["xy"], c5["z"], cb, cy,
                          geometric=True)
    eb = rel_l2(bp, rbp)
    print(f"C5 partial cache ({k} of {n_sub} subsets, {cb.size} bins): "
          f"forward rel L2 {ef:.2e}, max over significant bins {em:.2e}; "
          f"backprojection rel L2 {eb:.2e}", flush=True)
    assert ef < 2e-5 and em < 1e-3 and eb < 2e-5
