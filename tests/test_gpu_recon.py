"""Batched, graph-captured OSEM (fir_amd.recon) against the single-image
path and the reference's OSEM output."""
import numpy as np
import pytest

from conftest import rel_l2
from fir_amd import interfile as io
from fir_amd import lib as firlib
from fir_amd import recon
from fir_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_group_matches_reference_osem(c1_golden):
    """Three copies of the reference's run A in one group: every image
    equals the reference's 2 x 4 OSEM result."""
    w = synth.c1()
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    y = torch.from_numpy(c1_golden["y"]).cuda()
    g = recon.OSEMGroup(ctx, v, 4, [y, y.clone(), y.clone()])
    g.iteration_eager()
    g.iteration_eager()
    for x in g.x:
        assert rel_l2(x.cpu().numpy(), c1_golden["xa"]) < 1e-4
    ctx.close()


def test_graph_replays_match_eager(c1_golden):
    """Captured iterations replayed three times, with eager work and
    synchronisations between replays, track an eager twin after every
    replay."""
    w = synth.c1()
    v = firlib.VolDesc.from_spec(w.vol)
    rng = np.random.default_rng(21)
    ys = [torch.from_numpy(rng.poisson(c1_golden["y0"] * s).astype(np.float32))
          .cuda() for s in (0.5, 2.0, 8.0)]
    fw = (4.0, 4.0, 4.0)
    ctx_g = firlib.context_for(w)
    ctx_e = firlib.context_for(w)
    graphed = recon.OSEMGroup(ctx_g, v, 4, ys, fwhm=fw, conv_interval=1,
                              cut=55.0)
    eager = recon.OSEMGroup(ctx_e, v, 4, [y.clone() for y in ys], fwhm=fw,
                            conv_interval=1, cut=55.0)
    graphed.capture()          # runs iteration 1 eagerly, then captures
    eager.iteration_eager()
    for rep in range(3):
        graphed.iteration()
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work
        torch.cuda.synchronize()
        eager.iteration_eager()
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        for a, b in zip(graphed.x, eager.x):
            assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-5, rep
    ctx_g.close()
    ctx_e.close()


def test_batched_uneven_groups():
    """Solves of two image grids and subset counts 8, 16 and 24 (24 does
    not divide the 64 views: interleaved uneven subsets) in one
    BatchedOSEM; each equals the same solve run alone."""
    w = synth.c1()
    vols = [firlib.VolDesc.from_spec(w.vol),
            firlib.VolDesc.from_spec(io.VolSpec((48, 48, 8), (2.5, 2.5, 2.0),
                                                (-58.75, -58.75, 0.0)))]
    ctx = firlib.context_for(w)
    act = synth.phantom(w)[0].ravel()
    y0 = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(vols[0], torch.from_numpy(act).cuda(), y0)
    g = torch.Generator(device="cuda").manual_seed(7)
    solves = []
    for i, (vi, s) in enumerate([(0, 8), (1, 16), (0, 24), (1, 8), (0, 8)]):
        y = torch.poisson(y0 * (0.5 + i), generator=g)
        solves.append(dict(vol=vols[vi], n_subsets=s, y=y))
    b = recon.BatchedOSEM(w, solves)
    assert len(b.groups) == 4
    b.capture()
    b.iteration()
    out = [x.cpu().numpy() for x in b.images()]
    for i, s in enumerate(solves):
        c = firlib.context_for(w)
        c.set_uneven_subsets(True)
        single = recon.OSEMGroup(c, s["vol"], s["n_subsets"], [s["y"]])
        single.iteration_eager()
        single.iteration_eager()
        assert rel_l2(out[i], single.x[0].cpu().numpy()) < 1e-5, i
        c.close()
    ctx.close()


def test_batched_graph_with_gaussian_on_mixed_grids():
    """One graph for all groups, with the Gaussian every sub-iteration on
    grids of different voxel sizes (different taps) and sizes (the smaller
    grid first, so the larger one regrows the context's scratch after the
    smaller one's eager iteration): three replays later every solve still
    equals the same solve run alone."""
    w = synth.c1()
    vols = [firlib.VolDesc.from_spec(io.VolSpec((48, 48, 8), (2.5, 2.5, 2.0),
                                                (-58.75, -58.75, 0.0))),
            firlib.VolDesc.from_spec(w.vol)]
    ctx = firlib.context_for(w)
    act = synth.phantom(w)[0].ravel()
    y0 = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(vols[1], torch.from_numpy(act).cuda(), y0)
    g = torch.Generator(device="cuda").manual_seed(13)
    solves = []
    for i, (vi, s) in enumerate([(0, 8), (1, 16), (0, 16), (1, 8)]):
        y = torch.poisson(y0 * (0.5 + i), generator=g)
        solves.append(dict(vol=vols[vi], n_subsets=s, y=y))
    kw = dict(fwhm=(4.0, 4.0, 4.0), conv_interval=1, cut=50.0)
    b = recon.BatchedOSEM(w, solves, **kw)
    b.capture()
    for _ in range(3):
        b.iteration()
        torch.cuda.synchronize()
    out = [x.cpu().numpy() for x in b.images()]
    for i, s in enumerate(solves):
        c = firlib.context_for(w)
        single = recon.OSEMGroup(c, s["vol"], s["n_subsets"], [s["y"]], **kw)
        for _ in range(4):
            single.iteration_eager()
        assert rel_l2(out[i], single.x[0].cpu().numpy()) < 1e-5, i
        c.close()
    b.close()
    ctx.close()


def test_batched_uneven_groups_cached():
    """The same uneven solves on path caches (one shared context, caches of
    one grid sharing their path matrix). Each batched solve equals the same
    solve run alone on its cache (the per-image arithmetic is the same), and
    agrees with the plain solve: within the OSEM bar for 8 and 16 subsets;
    with 24 uneven subsets of 2-3 views each, 48 sub-iterations amplify the
    16-bit length error (about 2e-6 per projection) to about 2e-4. The
    2.5 mm grid is walked with the reference's int-truncated steps, which
    give some LORs negative lengths; the signed 16-bit lengths keep them."""
    w = synth.c1()
    vols = [firlib.VolDesc.from_spec(w.vol),
            firlib.VolDesc.from_spec(io.VolSpec((48, 48, 8), (2.5, 2.5, 2.0),
                                                (-58.75, -58.75, 0.0)))]
    ctx = firlib.context_for(w)
    act = synth.phantom(w)[0].ravel()
    y0 = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(vols[0], torch.from_numpy(act).cuda(), y0)
    g = torch.Generator(device="cuda").manual_seed(11)
    solves = []
    for i, (vi, s) in enumerate([(0, 8), (1, 16), (0, 24), (1, 8), (0, 8)]):
        y = torch.poisson(y0 * (0.5 + i), generator=g)
        solves.append(dict(vol=vols[vi], n_subsets=s, y=y))
    b = recon.BatchedOSEM(w, solves, cache_format=firlib.FIR_LEN_SNORM16)
    assert len(b.groups) == 4 and len(b.caches) == 4
    b.capture()
    b.iteration()
    out = [x.cpu().numpy() for x in b.images()]
    for grp, members in zip(b.groups, b.members):
        for i in members:
            s = solves[i]
            alone = recon.OSEMGroup(b.ctx, s["vol"], s["n_subsets"], [s["y"]],
                                    cache=grp.cache)
            alone.iteration_eager()
            alone.iteration_eager()
            assert rel_l2(out[i], alone.x[0].cpu().numpy()) < 1e-6, i
            c = firlib.context_for(w)
            c.set_uneven_subsets(True)
            plain = recon.OSEMGroup(c, s["vol"], s["n_subsets"], [s["y"]])
            plain.iteration_eager()
            plain.iteration_eager()
            bar = 1e-3 if s["n_subsets"] == 24 else 1e-4
            assert rel_l2(out[i], plain.x[0].cpu().numpy()) < bar, i
            c.close()
    b.close()
    ctx.close()
