"""The path cache (fir_cache_*): Siddon paths kept in device memory, with a
voxel-major transpose for atomic-free backprojection.

Bars. The cache reproduces the walk's elements exactly (same element count
as the on-the-fly walk, every LOR that crosses the volume is a row); its
lengths are stored in the chosen format, so its projections are compared
with the reference's outputs at that format's precision:
  FP32 lengths:     forward and backprojection within 1e-5 relative;
  snorm16 lengths:  within 2e-5 relative L2 (per-LOR scale, 2^-16 steps);
  FP16 lengths:     within 2e-4 relative L2 (11-bit significand).
OSEM images after the reference's schedules are within 1e-4 relative L2 of
the reference for the FP32 and snorm16 formats. The cached forward and
backward projectors are exact adjoints of each other (same stored matrix):
<A x, y> = <x, A^T y> to FP32 summation accuracy.
"""
import numpy as np
import pytest

import pyoracle as O
from conftest import rel_l2, rel_max
from fir_amd import interfile as io
from fir_amd import lib as firlib
from fir_amd import recon
from fir_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FORMATS = [firlib.FIR_LEN_FP32, firlib.FIR_LEN_SNORM16, firlib.FIR_LEN_FP16]
FWD_TOL = {firlib.FIR_LEN_FP32: 1e-5, firlib.FIR_LEN_SNORM16: 2e-5,
           firlib.FIR_LEN_FP16: 2e-4}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def c1():
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    w = synth.c1()
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    caches = {}
    yield w, ctx, v, caches
    for c in caches.values():
        c.close()
    ctx.close()


def cache_of(c1, n_subsets, fmt):
    w, ctx, v, caches = c1
    key = (n_subsets, fmt)
    if key not in caches:
        caches[key] = ctx.path_cache(v, n_subsets, fmt)
    return caches[key]


def test_cache_holds_every_path_element(c1):
    """Element count equals the on-the-fly walk's; rows are exactly the
    LORs with a non-empty path (oracle path dump)."""
    w, ctx, v, _ = c1
    cache = cache_of(c1, 1, firlib.FIR_LEN_FP32)
    info = cache.info()
    assert info["elements"] == ctx.path_count(v)
    xy, z = O.scanner_tables(w.scanner)
    counts, idx, _ = O.path_dump(w.vol, w.proj, xy, z)
    assert info["elements"] == int(np.maximum(counts, 0).sum())
    assert info["rows"] == int((counts > 0).sum())
    assert info["device_bytes"] >= info["elements"] * 16


@pytest.mark.parametrize("fmt", FORMATS)
def test_cached_forward_vs_reference(c1, c1_golden, fmt):
    w, ctx, v, _ = c1
    cache = cache_of(c1, 4, fmt)
    sino = torch.full((ctx.n_bins,), -1.0, device="cuda")
    cache.forward(dev(c1_golden["act"].ravel()), sino)
    y = host(sino)
    ref = c1_golden["y0"]
    assert np.all(y[ref == 0] == 0)          # misses and empty paths give 0
    assert rel_l2(y, ref) < FWD_TOL[fmt]
    if fmt == firlib.FIR_LEN_FP32:
        assert rel_max(y, ref) < 1e-5


@pytest.mark.parametrize("fmt", FORMATS)
def test_cached_backward_vs_reference(c1, c1_golden, fmt):
    w, ctx, v, _ = c1
    cache = cache_of(c1, 1, fmt)
    bp = torch.full((v.n_voxels,), -1.0, device="cuda")
    cache.backward(dev(c1_golden["y"]), bp)
    assert rel_l2(host(bp), c1_golden["bp"]) < FWD_TOL[fmt]
    if fmt == firlib.FIR_LEN_FP32:
        assert rel_max(host(bp), c1_golden["bp"]) < 1e-5


@pytest.mark.parametrize("fmt", [firlib.FIR_LEN_FP32, firlib.FIR_LEN_SNORM16])
def test_cached_sensitivity_vs_reference(c1, c1_golden, fmt):
    w, ctx, v, _ = c1
    cache = cache_of(c1, 4, fmt)
    sens = torch.zeros(4 * v.n_voxels, device="cuda")
    cache.sensitivity(sens)
    assert rel_max(host(sens), c1_golden["sens"]) < FWD_TOL[fmt]


@pytest.mark.parametrize("fmt", FORMATS)
def test_cached_pair_is_adjoint(c1, fmt):
    """<A x, y> == <x, A^T y> in FP64 sums of the FP32 results."""
    w, ctx, v, _ = c1
    cache = cache_of(c1, 1, fmt)
    g = torch.Generator().manual_seed(3)
    x = torch.rand(v.n_voxels, generator=g).cuda()
    y = torch.rand(ctx.n_bins, generator=g).cuda()
    ax = torch.zeros(ctx.n_bins, device="cuda")
    aty = torch.zeros(v.n_voxels, device="cuda")
    cache.forward(x, ax)
    cache.backward(y, aty)
    lhs = float((ax.double() * y.double()).sum())
    rhs = float((x.double() * aty.double()).sum())
    assert abs(lhs - rhs) / abs(lhs) < 1e-6


def test_cached_backward_matches_atomic_path(c1, c1_golden):
    """Gather backprojection over the transpose == the on-the-fly atomic
    scatter, per subset (FP32 lengths)."""
    w, ctx, v, _ = c1
    cache = cache_of(c1, 4, firlib.FIR_LEN_FP32)
    y = dev(c1_golden["y"])
    a = torch.zeros(4 * v.n_voxels, device="cuda")
    b = torch.zeros(4 * v.n_voxels, device="cuda")
    cache.backward(y, a)
    ctx.backward(v, y, b, 4)
    assert rel_max(host(a), host(b)) < 1e-5


@pytest.mark.parametrize("fmt", [firlib.FIR_LEN_FP32, firlib.FIR_LEN_SNORM16])
def test_cached_osem_vs_reference(c1, c1_golden, fmt):
    """Runs A (2 x 4), B (ACF, Gaussian every sub-iteration, cut 55 mm) and
    C (bias, 1 x 4) of the reference with the cached projector."""
    w, ctx, v, _ = c1
    g = c1_golden
    cache = cache_of(c1, 4, fmt)
    y = dev(g["y"])
    ga = recon.OSEMGroup(ctx, v, 4, [y], cache=cache)
    ga.iteration_eager()
    ga.iteration_eager()
    assert rel_l2(host(ga.x[0]), g["xa"]) < 1e-4

    mu = dev(g["hu"].ravel())
    ctx.hu_to_mu(mu)
    ctx.cut_circle(v, mu, 55.0)
    acf = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v, mu, acf)
    ctx.exp_clamped(acf)
    yb = dev(g["y"])
    ctx.mul_clamped(yb, acf)
    gb = recon.OSEMGroup(ctx, v, 4, [yb], fwhm=(4.0, 4.0, 4.0),
                         conv_interval=1, cut=55.0, cache=cache)
    gb.iteration_eager()
    gb.iteration_eager()
    assert rel_l2(host(gb.x[0]), g["xb"]) < 1e-4

    bias = torch.full((ctx.n_bins,), float(g["bias_value"]), device="cuda")
    gc = recon.OSEMGroup(ctx, v, 4, [y], biases=[bias], cache=cache)
    gc.iteration_eager()
    assert rel_l2(host(gc.x[0]), g["xc"]) < 1e-4


def test_family_path_osem_vs_reference(c1, c1_golden):
    """Runs A (2 x 4), B (ACF, Gaussian every sub-iteration, cut 55 mm) and
    C (bias) of the reference through the one-image family path (sheet-form
    line integrals, family-list backprojection), within 1e-4 of the
    reference's images."""
    w, ctx, v, _ = c1
    g = c1_golden
    cache = ctx.path_cache(v, 4, firlib.FIR_LEN_SNORM16,
                           row_order=firlib.CACHE_ROW_ORDER)
    y = dev(g["y"])
    # the family path is the one taken (its sums differ in order from the
    # transposed lists')
    xa = dev(g["act"].ravel() + 0.1)
    fam = torch.zeros(v.n_voxels, device="cuda")
    gat = torch.zeros(v.n_voxels, device="cuda")
    cache.osem_project([xa], [y], None, [fam], 0)
    cache.tune(8, 16, firlib.FIR_CACHE_NO_FAMILY_BP)
    cache.osem_project([xa], [y], None, [gat], 0)
    cache.tune(8, 16, 0)
    assert not torch.equal(fam, gat)
    ga = recon.OSEMGroup(ctx, v, 4, [y], cache=cache)
    ga.iteration_eager()
    ga.iteration_eager()
    assert rel_l2(host(ga.x[0]), g["xa"]) < 1e-4
    mu = dev(g["hu"].ravel())
    ctx.hu_to_mu(mu)
    ctx.cut_circle(v, mu, 55.0)
    acf = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v, mu, acf)
    ctx.exp_clamped(acf)
    yb = dev(g["y"])
    ctx.mul_clamped(yb, acf)
    gb = recon.OSEMGroup(ctx, v, 4, [yb], fwhm=(4.0, 4.0, 4.0),
                         conv_interval=1, cut=55.0, cache=cache)
    gb.iteration_eager()
    gb.iteration_eager()
    assert rel_l2(host(gb.x[0]), g["xb"]) < 1e-4
    bias = torch.full((ctx.n_bins,), float(g["bias_value"]), device="cuda")
    gc = recon.OSEMGroup(ctx, v, 4, [y], biases=[bias], cache=cache)
    gc.iteration_eager()
    assert rel_l2(host(gc.x[0]), g["xc"]) < 1e-4
    cache.close()


@pytest.mark.parametrize("n_images", [2, 3, 5, 8, 11])
def test_cached_batch_equals_single(c1, c1_golden, n_images):
    """Batched projector passes (interleaved images, 2/4/8 per matrix read)
    give each image exactly what a single-image pass gives."""
    w, ctx, v, _ = c1
    cache = cache_of(c1, 4, firlib.FIR_LEN_SNORM16)
    rng = np.random.default_rng(n_images)
    xs = [dev(rng.uniform(0.5, 2.0, v.n_voxels).astype(np.float32))
          for _ in range(n_images)]
    ys = [dev(rng.poisson(c1_golden["y0"] * s).astype(np.float32))
          for s in rng.uniform(0.1, 4.0, n_images)]
    bps = [torch.zeros(v.n_voxels, device="cuda") for _ in range(n_images)]
    cache.osem_project(xs, ys, None, bps, 1)
    for x, y, bp in zip(xs, ys, bps):
        one = torch.zeros(v.n_voxels, device="cuda")
        cache.osem_project([x], [y], None, [one], 1)
        assert np.array_equal(host(bp), host(one))


def test_cached_graph_replays_match_eager(c1, c1_golden):
    """A captured cached iteration replays identically to eager ones."""
    w, ctx, v, _ = c1
    cache = cache_of(c1, 4, firlib.FIR_LEN_SNORM16)
    rng = np.random.default_rng(5)
    ys = [dev(rng.poisson(c1_golden["y0"] * s).astype(np.float32))
          for s in (0.5, 3.0)]
    fw = (4.0, 4.0, 4.0)
    gg = recon.OSEMGroup(ctx, v, 4, ys, fwhm=fw, conv_interval=2, cut=55.0,
                         cache=cache)
    ge = recon.OSEMGroup(ctx, v, 4, [y.clone() for y in ys], fwhm=fw,
                         conv_interval=2, cut=55.0, cache=cache)
    gg.capture()
    ge.iteration_eager()
    for rep in range(3):
        gg.iteration()
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        ge.iteration_eager()
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        for a, b in zip(gg.x, ge.x):
            assert np.array_equal(host(a), host(b)), rep


def test_cache_layout_lanes_keep_results(c1, c1_golden):
    """Caches laid out for other lanes (fir_set_cache_lanes) and z-slice
    interleaves (fir_set_cache_zslices; 16 exceeds the 8 slices of the
    image, so the layout pads them), run with their own or other runtime
    lanes, give the same projections."""
    w, ctx, v, _ = c1
    ref = cache_of(c1, 4, firlib.FIR_LEN_FP32)
    img = dev(c1_golden["act"].ravel())
    y = dev(c1_golden["y"])
    a = torch.zeros(ctx.n_bins, device="cuda")
    ref.forward(img, a)
    fa = torch.zeros(4 * v.n_voxels, device="cuda")
    ref.backward(y, fa)
    for fl, bl, zq in [(4, 64, 1), (32, 2, 2), (1, 1, 8), (8, 16, 16)]:
        ctx.set_cache_lanes(fl, bl)
        ctx.set_cache_zslices(zq)
        try:
            c = ctx.path_cache(v, 4, firlib.FIR_LEN_FP32)
        finally:
            ctx.set_cache_lanes(8, 16)
            ctx.set_cache_zslices(4)
        for rl in [(fl, bl), (8, 16)]:
            c.tune(*rl)
            b = torch.zeros(ctx.n_bins, device="cuda")
            c.forward(img, b)
            fb = torch.zeros(4 * v.n_voxels, device="cuda")
            c.backward(y, fb)
            assert rel_max(host(b), host(a)) < 1e-6
            assert rel_max(host(fb), host(fa)) < 1e-6
        c.close()
    with pytest.raises(firlib.FirError):
        ctx.set_cache_lanes(3, 16)
    with pytest.raises(firlib.FirError):
        ctx.set_cache_zslices(3)


def test_cache_lanes_keep_results(c1, c1_golden):
    w, ctx, v, _ = c1
    cache = cache_of(c1, 4, firlib.FIR_LEN_FP32)
    img = dev(c1_golden["act"].ravel())
    outs = []
    # flags: XCD map (1), voxels walked in storage order (2)
    for fl, bl, fg in [(1, 1, 0), (4, 64, 1), (16, 16, 2), (64, 8, 3),
                       (8, 16, 0)]:
        cache.tune(fl, bl, fg)
        s = torch.zeros(ctx.n_bins, device="cuda")
        cache.forward(img, s)
        b = torch.zeros(4 * v.n_voxels, device="cuda")
        cache.backward(dev(c1_golden["y"]), b)
        outs.append((host(s), host(b)))
    cache.tune()
    for s, b in outs[1:]:
        assert rel_max(s, outs[0][0]) < 1e-6
        assert rel_max(b, outs[0][1]) < 1e-6
    for bad in [(0, 16, 0), (16, 3, 0), (128, 16, 0), (16, 16, 32)]:
        with pytest.raises(firlib.FirError):
            cache.tune(*bad)


def test_cache_keeps_negative_lengths():
    """The reference walk with int-truncated steps (2.5 mm voxels walked in
    2 mm steps) yields some negative lengths; the signed 16-bit cache keeps
    them: cached projections agree with the plain path (which reproduces
    them bit for bit) within the format's error."""
    w = synth.c1()
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(io.VolSpec((48, 48, 8), (2.5, 2.5, 2.0),
                                            (-58.75, -58.75, 0.0)))
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(v.n_voxels, generator=g, device="cuda")
    y = torch.rand(ctx.n_bins, generator=g, device="cuda")
    a = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v, x, a)
    bp = torch.zeros(v.n_voxels, device="cuda")
    ctx.backward(v, y, bp, 1)
    for fmt, tol in [(firlib.FIR_LEN_SNORM16, 1e-4), (firlib.FIR_LEN_FP32, 1e-6)]:
        cache = ctx.path_cache(v, 1, fmt)
        ca = torch.zeros(ctx.n_bins, device="cuda")
        cache.forward(x, ca)
        cb = torch.zeros(v.n_voxels, device="cuda")
        cache.backward(y, cb)
        assert rel_l2(host(ca), host(a)) < tol
        assert rel_l2(host(cb), host(bp)) < tol
        cache.close()
    ctx.close()


def test_cache_rejects_bad_arguments(c1):
    w, ctx, v, _ = c1
    with pytest.raises(firlib.FirError):
        ctx.path_cache(v, 5)                 # 5 does not divide 64 views
    with pytest.raises(firlib.FirError):
        ctx.path_cache(v, 4, 7)              # unknown length format
    cache = cache_of(c1, 4, firlib.FIR_LEN_SNORM16)
    x = torch.ones(v.n_voxels, device="cuda")
    y = torch.ones(ctx.n_bins, device="cuda")
    with pytest.raises(firlib.FirError):
        cache.osem_project([x], [y], None, [x.clone()], 4)


def test_c2_cache_properties():
    """Config-2 scale: the cache holds every element of the 3.9e9-element
    system matrix; cached forward agrees with the on-the-fly forward and
    the cached pair is adjoint at full size."""
    w = synth.c2()
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    cache = ctx.path_cache(v, 16, firlib.FIR_LEN_SNORM16)
    info = cache.info()
    assert info["elements"] == ctx.path_count(v)
    act, _ = synth.phantom(w)
    x = dev(act.ravel())
    a = torch.zeros(ctx.n_bins, device="cuda")
    b = torch.zeros(ctx.n_bins, device="cuda")
    cache.forward(x, a)
    ctx.forward(v, x, b)
    assert rel_l2(host(a), host(b)) < 2e-5
    g = torch.Generator(device="cuda").manual_seed(9)
    y = torch.rand(ctx.n_bins, device="cuda", generator=g)
    aty = torch.zeros(16 * v.n_voxels, device="cuda")
    cache.backward(y, aty)
    aty = aty.view(16, -1).sum(0)
    lhs = float((a.double() * y.double()).sum())
    rhs = float((x.double() * aty.double()).sum())
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    cache.close()
    # the sheet form at full size (rows in tangential-major order): one-image
    # OSEM passes equal the image gathers bit for bit (with the transposed
    # lists' backprojection), the family backprojection agrees with that
    # within FP32 summation order and repeats exactly, and both sum to the
    # on-the-fly projector's within the snorm16 bar
    cache = ctx.path_cache(v, 16, firlib.FIR_LEN_SNORM16,
                           row_order=firlib.CACHE_ROW_ORDER)
    yc = torch.poisson(b * 3.0, generator=g)
    for s in (0, 7):
        p1 = torch.zeros(v.n_voxels, device="cuda")
        cache.tune(8, 16, firlib.FIR_CACHE_NO_FAMILY_BP)
        cache.osem_project([x + 0.1], [yc], None, [p1], s)
        p2 = torch.zeros(v.n_voxels, device="cuda")
        cache.tune(8, 16, firlib.FIR_CACHE_NO_SHEETS)
        cache.osem_project([x + 0.1], [yc], None, [p2], s)
        assert torch.equal(p1, p2)
        pf = torch.full((v.n_voxels,), -1.0, device="cuda")
        pf2 = torch.zeros(v.n_voxels, device="cuda")
        cache.tune(8, 16, 0)
        cache.osem_project([x + 0.1], [yc], None, [pf], s)
        cache.osem_project([x + 0.1], [yc], None, [pf2], s)
        assert torch.equal(pf, pf2)
        assert float((pf - p1).abs().max()) <= 1e-6 * float(p1.abs().max())
        p3 = torch.zeros(v.n_voxels, device="cuda")
        ctx.osem_project(v, x + 0.1, yc, None, p3, s, 16)
        assert rel_l2(host(p1), host(p3)) < 2e-5
        assert rel_l2(host(pf), host(p3)) < 2e-5
    cache.close()
    ctx.close()


def test_autotuner_picks_fastest_and_keeps_results(c1, c1_golden):
    """The autotuner returns the fastest of its trials, tuned, and the
    chosen cache still reproduces the reference's OSEM run A."""
    from fir_amd import autotune
    w, ctx, v, _ = c1
    y = dev(c1_golden["y"])
    x = torch.ones(v.n_voxels, device="cuda")
    layouts = ((firlib.CACHE_ROW_ORDER, 16, 16, 1),
               (firlib.CACHE_ROW_ORDER, 8, 8, 4),
               (firlib.FIR_ORDER_VIEW_MAJOR, 8, 16, 2))
    cache, res = autotune.tune_cache(ctx, v, 4, [x], [y], reps=1,
                                     layouts=layouts)
    assert len(res.trials) == len(layouts)
    assert ctx._cache_lanes == (8, 16) and ctx._cache_zq == 4
    assert res.best_ms == min(t.ms for t in res.trials)
    g = recon.OSEMGroup(ctx, v, 4, [y], cache=cache)
    g.iteration_eager()
    g.iteration_eager()
    assert rel_l2(host(g.x[0]), c1_golden["xa"]) < 1e-4
    cache.close()
    pres = autotune.tune_plain(ctx, v, 4, x, y, reps=1)
    assert pres.best_ms == min(t.ms for t in pres.trials)
    ctx.set_lor_order(0)
    ctx.set_scatter_copies(1)


@pytest.mark.parametrize("fmt", FORMATS)
def test_sheet_form_equals_image_gathers(c1, c1_golden, fmt):
    """One-image OSEM passes over the sheet form (each (view, tangential
    bin) family's pixels x slices staged in LDS, 16-bit sheet indices) give
    the same results bit for bit as the gathers from the image
    (FIR_CACHE_NO_SHEETS): same elements, same order, same values (both
    backprojecting from the transposed lists)."""
    w, ctx, v, caches = c1
    cache = ctx.path_cache(v, 4, fmt, row_order=firlib.CACHE_ROW_ORDER)
    # the sheet form exists for the 16-bit length formats only
    assert cache.forms()["sheets"] == (fmt != firlib.FIR_LEN_FP32)
    x = dev(c1_golden["act"].ravel() + 0.1)
    y = dev(c1_golden["y"])
    bias = dev(np.full_like(c1_golden["y"], 0.5))
    for s in range(4):
        for b in (None, [bias]):
            a = torch.zeros(v.n_voxels, device="cuda")
            cache.tune(8, 16, firlib.FIR_CACHE_NO_FAMILY_BP)
            cache.osem_project([x], [y], b, [a], s)
            g = torch.zeros(v.n_voxels, device="cuda")
            cache.tune(8, 16, firlib.FIR_CACHE_NO_SHEETS)
            cache.osem_project([x], [y], b, [g], s)
            assert torch.equal(a, g) and float(a.abs().max()) > 0
    cache.close()


def test_cache_without_sheet_form(c1, c1_golden):
    """fir_set_cache_sheets(0): the matrix has no sheet indices (less device
    memory) and one-image passes gather from the image, with the same
    results."""
    w, ctx, v, caches = c1
    with_s = ctx.path_cache(v, 4, firlib.FIR_LEN_SNORM16,
                            row_order=firlib.CACHE_ROW_ORDER)
    ctx.set_cache_sheets(False)
    try:
        without = ctx.path_cache(v, 4, firlib.FIR_LEN_SNORM16,
                                 row_order=firlib.CACHE_ROW_ORDER)
    finally:
        ctx.set_cache_sheets(True)
    assert without.info()["device_bytes"] < with_s.info()["device_bytes"]
    x = dev(c1_golden["act"].ravel() + 0.1)
    y = dev(c1_golden["y"])
    a = torch.zeros(v.n_voxels, device="cuda")
    b = torch.zeros(v.n_voxels, device="cuda")
    with_s.tune(8, 16, firlib.FIR_CACHE_NO_FAMILY_BP)
    with_s.osem_project([x], [y], None, [a], 1)
    without.osem_project([x], [y], None, [b], 1)
    assert torch.equal(a, b)
    without.close()
    with_s.close()


@pytest.mark.parametrize("fmt", [firlib.FIR_LEN_SNORM16, firlib.FIR_LEN_FP16])
def test_family_backprojection(c1, c1_golden, fmt):
    """One-image OSEM passes over the sheet form backproject family by
    family (each (view, tangential bin) family's per-voxel element lists
    summed against its weights in LDS, then the families' partial sheets
    summed per voxel): within FP32 summation order of the transposed lists'
    backprojection (1e-6 of the largest value), the same bits on every run,
    every voxel written, and within the format's bar of the on-the-fly
    projector."""
    w, ctx, v, caches = c1
    cache = ctx.path_cache(v, 4, fmt, row_order=firlib.CACHE_ROW_ORDER)
    assert cache.forms() == {"sheets": True, "family_lists": True,
                             "slabs": True}
    x = dev(c1_golden["act"].ravel() + 0.1)
    y = dev(c1_golden["y"])
    bias = dev(np.full_like(c1_golden["y"], 0.5))
    for s in range(4):
        for b in (None, [bias]):
            f1 = torch.full((v.n_voxels,), -7.0, device="cuda")
            f2 = torch.full((v.n_voxels,), 3.0, device="cuda")
            cache.tune(8, 16, 0)
            cache.osem_project([x], [y], b, [f1], s)
            cache.osem_project([x], [y], b, [f2], s)
            assert torch.equal(f1, f2) and float(f1.min()) >= 0
            g = torch.zeros(v.n_voxels, device="cuda")
            cache.tune(8, 16, firlib.FIR_CACHE_NO_FAMILY_BP)
            cache.osem_project([x], [y], b, [g], s)
            assert float((f1 - g).abs().max()) <= 1e-6 * float(g.abs().max())
            p = torch.zeros(v.n_voxels, device="cuda")
            ctx.osem_project(v, x, y, b[0] if b else None, p, s, 4)
            assert rel_l2(host(f1), host(p)) < FWD_TOL[fmt]
    cache.tune(8, 16, 0)
    cache.close()


@pytest.mark.parametrize("n_images", [2, 3, 8, 11])
def test_family_batch_backprojection(c1, c1_golden, n_images):
    """Batches over a cache with family lists take their line integrals
    from slab-staged sheets (4 or 8 images; 2 gather from the image copies)
    and backproject all their images family by family from the lists (one
    read of the lists per batch, weights of every image in LDS): each image
    gets exactly what its one-image family pass gives (the same weights,
    lists and summation order), and exactly what the gathered line
    integrals give; within FP32 summation order of the transposed-list
    backprojection (which the cache builds only when tuning asks for it),
    and the full backprojection and sensitivity through the lists match the
    transposed ones the same way."""
    w, ctx, v, caches = c1
    cache = ctx.path_cache(v, 4, firlib.FIR_LEN_SNORM16,
                           row_order=firlib.CACHE_ROW_ORDER)
    forms = cache.forms()
    assert forms["family_lists"] and forms["slabs"], forms
    rng = np.random.default_rng(40 + n_images)
    xs = [dev(rng.uniform(0.5, 2.0, v.n_voxels).astype(np.float32))
          for _ in range(n_images)]
    ys = [dev(rng.poisson(c1_golden["y0"] * s).astype(np.float32))
          for s in rng.uniform(0.1, 4.0, n_images)]
    bps = [torch.full((v.n_voxels,), -1.0, device="cuda")
           for _ in range(n_images)]
    bytes_before = cache.info()["device_bytes"]
    cache.osem_project(xs, ys, None, bps, 2)
    sens_f = torch.zeros(4 * v.n_voxels, device="cuda")
    cache.sensitivity(sens_f)
    back_f = torch.zeros(4 * v.n_voxels, device="cuda")
    cache.backward(ys[0], back_f)
    for x, y, bp in zip(xs, ys, bps):
        one = torch.zeros(v.n_voxels, device="cuda")
        cache.osem_project([x], [y], None, [one], 2)
        assert np.array_equal(host(bp), host(one))
    cache.tune(8, 16, firlib.FIR_CACHE_NO_SLABS)
    gathered = [torch.zeros(v.n_voxels, device="cuda") for _ in xs]
    cache.osem_project(xs, ys, None, gathered, 2)
    for bp, g in zip(bps, gathered):
        assert np.array_equal(host(bp), host(g))
    cache.tune(8, 16, firlib.FIR_CACHE_NO_FAMILY_BP)
    assert cache.info()["device_bytes"] > bytes_before  # transposes built
    for x, y, bp in zip(xs, ys, bps):
        g = torch.zeros(v.n_voxels, device="cuda")
        cache.osem_project([x], [y], None, [g], 2)
        assert float((bp - g).abs().max()) <= 1e-6 * float(g.abs().max())
    sens_t = torch.zeros(4 * v.n_voxels, device="cuda")
    cache.sensitivity(sens_t)
    back_t = torch.zeros(4 * v.n_voxels, device="cuda")
    cache.backward(ys[0], back_t)
    assert rel_max(host(sens_f), host(sens_t)) <= 1e-6
    assert rel_max(host(back_f), host(back_t)) <= 1e-6
    cache.close()
