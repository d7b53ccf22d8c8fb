"""Parity of the HIP path (through the C-ABI) with the reference's outputs
and the CPU oracle.

Bars: bit-exact for Siddon index sequences/lengths and the forward
projection (the walk and the FP32 accumulation reproduce the reference's
operation order); 1e-5 relative for backprojection and sensitivity (FP32
atomic reordering); 1e-4 relative L2 for OSEM images.
"""
import hashlib

import numpy as np
import pytest

import pyoracle as O
from conftest import rel_l2, rel_max
from fir_amd import interfile as io
from fir_amd import lib as firlib
from fir_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def c1ctx():
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    w = synth.c1()
    ctx = firlib.context_for(w)
    yield w, ctx
    ctx.close()


def gpu_paths(ctx, vol, n_bins, bin0=0):
    v = firlib.VolDesc.from_spec(vol) if not isinstance(
        vol, firlib.VolDesc) else vol
    maxp = v.nx + v.ny + v.nz
    counts = torch.zeros(n_bins, dtype=torch.int32, device="cuda")
    idx = torch.zeros(n_bins * maxp, dtype=torch.int32, device="cuda")
    ln = torch.zeros(n_bins * maxp, dtype=torch.float64, device="cuda")
    ctx.path_dump(v, bin0, n_bins, maxp, counts, idx, ln)
    c = host(counts)
    i2 = host(idx).reshape(n_bins, maxp)
    l2 = host(ln).reshape(n_bins, maxp)
    mask = np.arange(maxp)[None, :] < np.maximum(c, 0)[:, None]
    return c, i2[mask], l2[mask]


@pytest.mark.parametrize("tag", ["c1", "t1"])
def test_paths_bitexact_vs_reference(c1ctx, paths_golden, tag):
    w, ctx = c1ctx
    meta, arr = paths_golden
    m = meta[tag]
    vol = io.VolSpec(m["vol"]["size"], m["vol"]["voxel"], m["vol"]["offset"])
    counts, idx, ln = gpu_paths(ctx, vol, m["n_bins"])
    assert np.array_equal(counts, arr[f"{tag}_counts"].astype(np.int32))
    assert hashlib.sha256(idx.tobytes()).hexdigest() == m["idx_sha256"]
    assert hashlib.sha256(ln.tobytes()).hexdigest() == m["len_sha256"]


def test_forward_bitexact_vs_reference(c1ctx, c1_golden):
    w, ctx = c1ctx
    v = firlib.VolDesc.from_spec(w.vol)
    img = dev(c1_golden["act"].ravel())
    sino = torch.full((ctx.n_bins,), -1.0, device="cuda")
    ctx.forward(v, img, sino)
    assert np.array_equal(host(sino), c1_golden["y0"])


def test_forward_subsets_cover_all_bins(c1ctx, c1_golden):
    w, ctx = c1ctx
    v = firlib.VolDesc.from_spec(w.vol)
    img = dev(c1_golden["act"].ravel())
    sino = torch.full((ctx.n_bins,), -1.0, device="cuda")
    for s in range(4):
        ctx.forward(v, img, sino, subset=s, n_subsets=4)
    assert np.array_equal(host(sino), c1_golden["y0"])


def test_backward_vs_reference(c1ctx, c1_golden):
    w, ctx = c1ctx
    v = firlib.VolDesc.from_spec(w.vol)
    y = dev(c1_golden["y"])
    img = torch.full((v.n_voxels,), 7.0, device="cuda")
    ctx.backward(v, y, img, 1)
    bp = host(img)
    assert rel_max(bp, c1_golden["bp"]) < 1e-5
    xy, z = O.scanner_tables(w.scanner)
    bp64 = O.backward(w.vol, w.proj, xy, z, c1_golden["y"].astype(np.float64))
    assert rel_max(bp, bp64) < 1e-5


def test_sensitivity_vs_reference(c1ctx, c1_golden):
    w, ctx = c1ctx
    v = firlib.VolDesc.from_spec(w.vol)
    sens = torch.zeros(4 * v.n_voxels, device="cuda")
    ctx.sensitivity(v, sens, 4)
    assert rel_max(host(sens), c1_golden["sens"]) < 1e-5


def run_osem(ctx, w, y, sens, n_iter, n_sub, cut=0.0, conv_interval=0,
             fwhm=(0.0, 0.0, 0.0), bias=None):
    """reconAlgos::OSEM (reconAlgos.cc:63-201) driven through the C-ABI."""
    v = firlib.VolDesc.from_spec(w.vol)
    n = v.n_voxels
    x = torch.ones(n, device="cuda")
    bp = torch.zeros(n, device="cuda")
    ctx.cut_circle(v, x, cut)
    conv = conv_interval > 0 and all(f > 0 for f in fwhm)
    total = n_iter * n_sub
    for it in range(n_iter):
        for s in range(n_sub):
            sub = it * n_sub + s + 1
            ctx.osem_subiter(v, x, y, bias, sens[s * n:(s + 1) * n], bp, s,
                             n_sub, convolve=conv and sub % conv_interval == 0,
                             fwhm=fwhm, cut=cut, reset=sub != total)
    return host(x)


def test_osem_vs_reference(c1ctx, c1_golden):
    w, ctx = c1ctx
    g = c1_golden
    v = firlib.VolDesc.from_spec(w.vol)
    sens = torch.zeros(4 * v.n_voxels, device="cuda")
    ctx.sensitivity(v, sens, 4)
    y = dev(g["y"])
    xa = run_osem(ctx, w, y, sens, 2, 4)
    assert rel_l2(xa, g["xa"]) < 1e-4
    # ACF chain on the GPU: HU -> mu, cut, forward, exp, y *= ACF
    mu = dev(g["hu"].ravel())
    ctx.hu_to_mu(mu)
    ctx.cut_circle(v, mu, 55.0)
    acf = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v, mu, acf)
    ctx.exp_clamped(acf)
    yb = dev(g["y"])
    ctx.mul_clamped(yb, acf)
    xb = run_osem(ctx, w, yb, sens, 2, 4, cut=55.0, conv_interval=1,
                  fwhm=(4.0, 4.0, 4.0))
    assert rel_l2(xb, g["xb"]) < 1e-4
    bias = torch.full((ctx.n_bins,), float(g["bias_value"]), device="cuda")
    xc = run_osem(ctx, w, y, sens, 1, 4, bias=bias)
    assert rel_l2(xc, g["xc"]) < 1e-4


def test_acf_chain_vs_reference(c1ctx, c1_golden):
    w, ctx = c1ctx
    v = firlib.VolDesc.from_spec(w.vol)
    mu = dev(c1_golden["hu"].ravel())
    ctx.hu_to_mu(mu)
    assert np.array_equal(host(mu), O.hu_to_mu(c1_golden["hu"]))
    acf = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v, mu, acf)
    ctx.exp_clamped(acf)
    assert rel_max(host(acf), c1_golden["acf"]) < 1e-6


@pytest.mark.parametrize("workload", ["c2", "c1"])
def test_convolve_and_cut_vs_oracle(workload):
    """The Gaussian (edge restore and cut included) equals the oracle's
    FP32 convolution bit for bit: at C2 through the vectorised passes (four
    outputs per thread), at C1 through the one-voxel-per-thread ones; also
    with a z FWHM whose half-width differs from x and y's."""
    w = getattr(synth, workload)()
    ctx = firlib.Context(0)
    ctx.set_stream(torch.cuda.current_stream())
    v = firlib.VolDesc.from_spec(w.vol)
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 3, v.n_voxels).astype(np.float32)
    half = 0.5 * v.nx * v.dx
    for fwhm, cut in (((4.0, 4.0, 4.0), 0.0), ((4.0, 4.0, 4.0), 0.9 * half),
                      ((3.0, 5.0, 8.0), 0.8 * half)):
        t = dev(img)
        ctx.convolve(v, t, fwhm, cut)
        ref = O.convolve(w.vol, img, fwhm, cut)
        assert np.array_equal(host(t), ref), (fwhm, cut)
    t = dev(img)
    ctx.cut_circle(v, t, 120.0)
    assert np.array_equal(host(t), O.cut_circle(w.vol, img, 120.0))
    ctx.close()


def test_clamped_ops_vs_oracle():
    ctx = firlib.Context(0)
    ctx.set_stream(torch.cuda.current_stream())
    rng = np.random.default_rng(9)
    a = rng.uniform(-1e-4, 3e-4, 100000).astype(np.float32)
    b = rng.uniform(-1e-4, 3e-4, 100000).astype(np.float32)
    a[:4] = [1e-5, np.nextafter(np.float32(1e-5), 1), 0, -0.0]
    for op, ref in (("mul_clamped", O.mul_clamped),
                    ("div_clamped", O.div_clamped)):
        t = dev(a)
        getattr(ctx, op)(t, dev(b))
        assert np.array_equal(host(t), ref(a, b)), op
    e = rng.uniform(-1, 5, 100000).astype(np.float32)
    t = dev(e)
    ctx.exp_clamped(t)
    # GPU expf and glibc expf may differ in the last place
    np.testing.assert_allclose(host(t), O.exp_clamped(e), rtol=3e-7)
    hu = rng.uniform(-2000, 3000, 100000).astype(np.float32)
    t = dev(hu)
    ctx.hu_to_mu(t)
    assert np.array_equal(host(t), O.hu_to_mu(hu))
    ctx.close()


def test_c2_paths_and_forward_sample():
    """Full-size C2 geometry: a strided sample of bins has bit-identical
    paths and forward values to the oracle."""
    w = synth.c2()
    ctx = firlib.context_for(w)
    xy, z = O.scanner_tables(w.scanner)
    nb = ctx.n_bins
    rng = np.random.default_rng(1)
    starts = rng.integers(0, nb - 64, 40)
    img = rng.uniform(0, 2, w.vol.n_voxels).astype(np.float32)
    sino = torch.zeros(nb, device="cuda")
    ctx.forward(firlib.VolDesc.from_spec(w.vol), dev(img), sino)
    s = host(sino)
    for b0 in starts:
        b0 = int(b0)
        gc, gi, gl = gpu_paths(ctx, w.vol, 64, bin0=b0)
        oc, oi, ol = O.path_dump(w.vol, w.proj, xy, z, b0, 64)
        assert np.array_equal(gc, oc)
        assert np.array_equal(gi, oi)
        assert np.array_equal(gl, ol)
        # reference accumulation order on the oracle's path
        k = 0
        for j, c in enumerate(oc):
            acc = np.float32(0)
            for e in range(max(c, 0)):
                acc = np.float32(np.float64(acc) + ol[k + e] *
                                 np.float64(img[oi[k + e]]))
            k += max(c, 0)
            assert s[b0 + j] == acc
    ctx.close()


def test_c2_adjoint():
    """<A x, y> = <x, A^T y> at the full C2 size (FP32 kernels, FP64 dots)."""
    w = synth.c2()
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.rand(v.n_voxels, device="cuda", generator=g)
    y = torch.rand(ctx.n_bins, device="cuda", generator=g)
    ax = torch.zeros(ctx.n_bins, device="cuda")
    aty = torch.zeros(v.n_voxels, device="cuda")
    ctx.forward(v, x, ax)
    ctx.backward(v, y, aty, 1)
    torch.cuda.synchronize()
    lhs = float((ax.double() * y.double()).sum())
    rhs = float((x.double() * aty.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    ctx.close()


def test_c5_geometric_paths_sample():
    """Long-axial 64x672 geometry with 1.6 mm voxels in geometric mode: a
    sample of bins has bit-identical paths to the oracle."""
    w = synth.c5()
    ctx = firlib.context_for(w, geometric=True)
    xy, z = O.scanner_tables(w.scanner)
    rng = np.random.default_rng(2)
    for b0 in rng.integers(0, ctx.n_bins - 32, 12):
        b0 = int(b0)
        gc, gi, gl = gpu_paths(ctx, w.vol, 32, bin0=b0)
        oc, oi, ol = O.path_dump(w.vol, w.proj, xy, z, b0, 32,
                                 geometric=True)
        assert np.array_equal(gc, oc)
        assert np.array_equal(gi, oi) and np.array_equal(gl, ol)
    ctx.close()


def test_errors_are_reported():
    w = synth.c1()
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    x = torch.ones(v.n_voxels, device="cuda")
    bp = torch.zeros(v.n_voxels, device="cuda")
    y = torch.ones(ctx.n_bins, device="cuda")
    with pytest.raises(firlib.FirError, match="divisor of 64"):
        ctx.osem_project(v, x, y, None, bp, 0, 3)
    bad = firlib.VolDesc(8, 8, 8, 0.9, 2.0, 2.0, 0, 0, 0)
    with pytest.raises(firlib.FirError, match="truncates to 0"):
        ctx.forward(bad, x, y)
    with pytest.raises(firlib.FirError, match="subset 4 out of range"):
        ctx.backward_subset(v, y, bp, 4, 4)
    ctx.close()


def test_batched_projector_matches_single(c1ctx, c1_golden):
    """fir_osem_project_batch on 11 images (two launches: 8 + 3) gives each
    image what fir_osem_project gives it."""
    w, ctx = c1ctx
    v = firlib.VolDesc.from_spec(w.vol)
    rng = np.random.default_rng(12)
    n = 11
    xs = [dev(rng.uniform(0.2, 2, v.n_voxels).astype(np.float32))
          for _ in range(n)]
    ys = [dev(rng.poisson(c1_golden["y0"] * (i + 1) / 4).astype(np.float32))
          for i in range(n)]
    bias = [None if i % 3 else torch.full((ctx.n_bins,), 0.1, device="cuda")
            for i in range(n)]
    bps = [torch.zeros(v.n_voxels, device="cuda") for _ in range(n)]
    ctx.osem_project_batch(v, xs, ys, bias, bps, 1, 4)
    for i in range(n):
        ref = torch.zeros(v.n_voxels, device="cuda")
        ctx.osem_project(v, xs[i], ys[i], bias[i], ref, 1, 4)
        assert rel_max(host(bps[i]), host(ref)) < 1e-5, i


def test_uneven_subsets_partition_the_views(c1ctx, c1_golden):
    """With uneven subsets enabled, 5 subsets of the 64 views (interleaved,
    13/13/13/13/12 views) cover every bin exactly once, and their
    backprojections sum to the single-subset backprojection."""
    w, ctx = c1ctx
    v = firlib.VolDesc.from_spec(w.vol)
    img = dev(c1_golden["act"].ravel())
    with pytest.raises(firlib.FirError, match="divisor"):
        ctx.forward(v, img, torch.zeros(ctx.n_bins, device="cuda"),
                    subset=0, n_subsets=5)
    ctx.set_uneven_subsets(True)
    try:
        sino = torch.full((ctx.n_bins,), -1.0, device="cuda")
        for s in range(5):
            ctx.forward(v, img, sino, subset=s, n_subsets=5)
        assert np.array_equal(host(sino), c1_golden["y0"])
        y = dev(c1_golden["y"])
        parts = torch.zeros(5 * v.n_voxels, device="cuda")
        ctx.backward(v, y, parts, 5)
        total = host(parts).reshape(5, -1).sum(0)
        assert rel_max(total, c1_golden["bp"]) < 1e-5
    finally:
        ctx.set_uneven_subsets(False)


@pytest.mark.parametrize("order,copies", [(0x100, 1), (0x401, 8), (0x301, 4),
                                          (0x200, 2), (0x600, 1)])
def test_tuning_knobs_keep_results(c1ctx, c1_golden, order, copies):
    """Row order and per-XCD scatter copies change only FP32 summation
    order: forward stays bit-exact, backprojection within 1e-5."""
    w, ctx = c1ctx
    v = firlib.VolDesc.from_spec(w.vol)
    ctx.set_lor_order(order)
    ctx.set_scatter_copies(copies)
    try:
        img = dev(c1_golden["act"].ravel())
        sino = torch.full((ctx.n_bins,), -1.0, device="cuda")
        ctx.forward(v, img, sino)
        assert np.array_equal(host(sino), c1_golden["y0"])
        bp = torch.zeros(v.n_voxels, device="cuda")
        ctx.backward(v, dev(c1_golden["y"]), bp, 1)
        assert rel_max(host(bp), c1_golden["bp"]) < 1e-5
        x = torch.ones(v.n_voxels, device="cuda")
        a = torch.zeros(v.n_voxels, device="cuda")
        ctx.osem_project(v, x, dev(c1_golden["y"]), None, a, 2, 4)
        b2 = [torch.zeros(v.n_voxels, device="cuda") for _ in range(3)]
        ctx.osem_project_batch(v, [x, x, x], [dev(c1_golden["y"])] * 3, None,
                               b2, 2, 4)
        for t in b2:
            assert rel_max(host(t), host(a)) < 1e-5
    finally:
        ctx.set_lor_order(0)
        ctx.set_scatter_copies(1)


def test_lor_order_rejects_bad_flags(c1ctx):
    _, ctx = c1ctx
    for bad in (0x500, 1, 0x800, 2):
        with pytest.raises(firlib.FirError):
            ctx.set_lor_order(bad)
    ctx.set_lor_order(0)
