"""Partial path caches (fir_cache_create_partial): the paths of subsets
0 .. k-1 only, for scanners whose whole matrix exceeds device memory; the
other subsets are walked on the fly.

Bars. A partial cache holds exactly the cached subsets' paths (element
counts equal the on-the-fly walk's per subset), and its projections of
those subsets are bit-identical to a whole cache's: the rows of a cached
view are the same paths in the same order, so every sum runs in the same
order. Bins and frames of the other subsets are left untouched.
"""
import numpy as np
import pytest

import pyoracle as O
from fir_amd import lib as firlib
from fir_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

S = 4


@pytest.fixture(scope="module")
def c1(c1_golden):
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    w = synth.c1()
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    whole = ctx.path_cache(v, S)
    part = ctx.path_cache(v, S, max_subsets=2)
    yield w, ctx, v, whole, part, c1_golden
    part.close()
    whole.close()
    ctx.close()


def bin_views(spec):
    """View of every bin: segments in order, each [view][axial][tang]
    (ProjData.h:16-21)."""
    p = O.proj_desc(spec)
    v = np.concatenate([np.repeat(np.arange(p.n_views), na * p.n_tang)
                        for na in p.n_axial])
    assert v.size == p.n_bins
    return v


def test_partial_cache_holds_the_first_subsets(c1):
    w, ctx, v, whole, part, _ = c1
    assert whole.n_cached == S and part.n_cached == 2
    per_subset = [ctx.path_count(v, s, S) for s in range(S)]
    assert sum(per_subset) == whole.info()["elements"]
    assert part.info()["elements"] == per_subset[0] + per_subset[1]
    assert part.info()["device_bytes"] < whole.info()["device_bytes"]


def test_partial_osem_project_equals_whole_cache(c1):
    w, ctx, v, whole, part, g = c1
    x = torch.from_numpy(g["act"].ravel() + 0.1).cuda()
    y = torch.from_numpy(g["y"]).cuda()
    for s in range(2):
        a = torch.zeros(v.n_voxels, device="cuda")
        b = torch.full((v.n_voxels,), -7.0, device="cuda")
        whole.osem_project([x], [y], None, [a], s)
        part.osem_project([x], [y], None, [b], s)
        assert torch.equal(a, b) and float(a.abs().max()) > 0
    with pytest.raises(firlib.FirError, match="not in the path cache"):
        part.osem_project([x], [y], None, [a], 2)


def test_partial_forward_backward_leave_other_subsets(c1):
    w, ctx, v, whole, part, g = c1
    x = torch.from_numpy(g["act"].ravel()).cuda()
    ref = torch.zeros(ctx.n_bins, device="cuda")
    whole.forward(x, ref)
    got = torch.full((ctx.n_bins,), -7.0, device="cuda")
    part.forward(x, got)
    ref, got = ref.cpu().numpy(), got.cpu().numpy()
    cached = bin_views(w.proj) % S < 2
    assert np.array_equal(got[cached], ref[cached])
    assert (got[~cached] == -7.0).all()
    assert (ref[cached] == 0).any() and (ref[cached] > 0).any()

    y = torch.from_numpy(g["y"]).cuda()
    fa = torch.zeros(S * v.n_voxels, device="cuda")
    fb = torch.full((S * v.n_voxels,), -7.0, device="cuda")
    whole.backward(y, fa)
    part.backward(y, fb)
    fa = fa.view(S, -1).cpu().numpy()
    fb = fb.view(S, -1).cpu().numpy()
    assert np.array_equal(fb[:2], fa[:2])
    assert (fb[2:] == -7.0).all()


def test_partial_plus_walk_gives_whole_backprojection(c1):
    """Cached frames of subsets 0-1 plus the on-the-fly walk's frames of
    subsets 2-3 reproduce the one-frame backprojection (reference
    BackProj output) within 1e-5 relative."""
    w, ctx, v, whole, part, g = c1
    y = torch.from_numpy(g["y"]).cuda()
    frames = torch.zeros(S * v.n_voxels, device="cuda")
    for s in range(2, S):
        ctx.backward_subset(v, y, frames[s * v.n_voxels:(s + 1) * v.n_voxels],
                            s, S)
    part.backward(y, frames)
    total = frames.view(S, -1).sum(0).cpu().numpy().astype(np.float64)
    ref = g["bp"].astype(np.float64).ravel()
    assert np.linalg.norm(total - ref) / np.linalg.norm(ref) < 1e-5


def test_partial_cache_budget(c1):
    w, ctx, v, whole, part, _ = c1
    with pytest.raises(firlib.FirError, match="subset 0 alone needs"):
        ctx.path_cache(v, S, max_bytes=1)
    big = ctx.path_cache(v, S, max_bytes=1 << 40)
    assert big.n_cached == S
    assert big.info()["elements"] == whole.info()["elements"]
    big.close()


def test_chunked_transpose_gives_the_same_cache(c1):
    """A transpose in many small steps (voxel histogram first, then sorted
    voxel ranges) stores the same voxel lists as the one-step transpose
    (counts taken from the sorted keys): bit-identical projections."""
    w, ctx, v, whole, part, g = c1
    x = torch.from_numpy(g["act"].ravel() + 0.1).cuda()
    y = torch.from_numpy(g["y"]).cuda()
    ctx.set_cache_chunk(20_000)
    try:
        small = ctx.path_cache(v, S)
    finally:
        ctx.set_cache_chunk(1 << 28)
    assert small.info()["elements"] > 10 * 20_000
    fa = torch.zeros(S * v.n_voxels, device="cuda")
    fb = torch.zeros(S * v.n_voxels, device="cuda")
    whole.backward(y, fa)
    small.backward(y, fb)
    assert torch.equal(fa, fb)
    for s in range(S):
        a = torch.zeros(v.n_voxels, device="cuda")
        b = torch.zeros(v.n_voxels, device="cuda")
        whole.osem_project([x], [y], None, [a], s)
        small.osem_project([x], [y], None, [b], s)
        assert torch.equal(a, b)
    with pytest.raises(firlib.FirError):
        ctx.set_cache_chunk(10)
    small.close()
