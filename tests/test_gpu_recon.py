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


def _c1_uneven_solves(n):
    w = synth.c1()
    ctx = firlib.context_for(w)
    v64 = firlib.VolDesc.from_spec(w.vol)
    v48 = firlib.VolDesc.from_spec(io.VolSpec((48, 48, 8), (2.0, 2.0, 2.0),
                                              (-47.0, -47.0, 0.0)))
    act, _ = synth.phantom(w)
    y0 = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v64, torch.from_numpy(act.ravel()).cuda(), y0)
    y0 = y0 / y0.sum()
    solves = []
    for i in range(n):
        g = torch.Generator(device="cuda").manual_seed(100 + i)
        solves.append(dict(vol=v48 if i % 2 else v64,
                           n_subsets=(2, 4, 8)[i % 3],
                           y=torch.poisson(y0 * 1e5 * (1 + i), generator=g)))
    ctx.close()
    return w, solves


@pytest.mark.parametrize("conv", [0, 1])
def test_multi_step_equals_group_graphs_c1(conv):
    """fir_osem_step (MultiOSEM): 21 solves on two grids with 2, 4 and 8
    subsets -- jobs of 8 images and of the rest, 4-image slots for 4 or
    fewer -- get exactly what the per-group graphs give, with and without
    the Gaussian."""
    w, solves = _c1_uneven_solves(21)
    kw = dict(fwhm=w.fwhm, conv_interval=conv, cut=0.0 if conv else 60.0)
    b = recon.BatchedOSEM(w, solves, cache_format=firlib.FIR_LEN_SNORM16,
                          **kw)
    b.capture()
    b.iteration()
    torch.cuda.synchronize()
    ref = [x.cpu().numpy() for x in b.images()]
    b.close()
    m = recon.MultiOSEM(w, solves, **kw)
    m.capture()
    m.iteration()
    torch.cuda.synchronize()
    out = [x.cpu().numpy() for x in m.images()]
    m.close()
    assert all(np.isfinite(o).all() for o in out)
    diff = [i for i in range(len(out)) if not np.array_equal(out[i], ref[i])]
    assert not diff, diff


def test_multi_step_mixed_voxel_sizes_c1():
    """A step whose images do not share the Gaussian's half-widths (a grid
    of 4 mm pixels beside the 2 mm ones) runs the one-voxel-per-thread
    passes for all of them instead of the vectorised ones, and still gives
    exactly what the per-group graphs give."""
    w, solves = _c1_uneven_solves(8)
    v32 = firlib.VolDesc.from_spec(io.VolSpec((32, 32, 8), (4.0, 4.0, 2.0),
                                              (-62.0, -62.0, 0.0)))
    for i in (1, 3, 5):
        solves[i] = dict(solves[i], vol=v32)
    kw = dict(fwhm=w.fwhm, conv_interval=1)
    b = recon.BatchedOSEM(w, solves, cache_format=firlib.FIR_LEN_SNORM16,
                          **kw)
    b.capture()
    b.iteration()
    torch.cuda.synchronize()
    ref = [x.cpu().numpy() for x in b.images()]
    b.close()
    m = recon.MultiOSEM(w, solves, **kw)
    m.capture()
    m.iteration()
    torch.cuda.synchronize()
    out = [x.cpu().numpy() for x in m.images()]
    m.close()
    assert all(np.isfinite(o).all() for o in out)
    diff = [i for i in range(len(out)) if not np.array_equal(out[i], ref[i])]
    assert not diff, diff


def test_multi_step_autotuned_layout_c1():
    """MultiOSEM with autotune_layout: the tuner times one step of the
    largest group's first job on a cache of each candidate layout (LOR sort
    key, lanes per LOR), keeps the fastest, and every solve then gets
    exactly what the per-group graphs give on caches of that layout."""
    w, solves = _c1_uneven_solves(21)
    kw = dict(fwhm=w.fwhm, conv_interval=1)
    m = recon.MultiOSEM(w, solves, autotune_layout=True, **kw)
    t = m.tuning
    assert len(t.trials) >= 3, t.table()
    assert t.best_ms == min(tr.ms for tr in t.trials)
    assert (m.cache_lanes, m.cache_row_order) == (
        t.best["fwd_lanes"], t.best["row_order"])
    lanes = {tr.config["fwd_lanes"] for tr in t.trials}
    assert {2, 4, 8} <= lanes, t.table()
    m.capture()
    m.iteration()
    torch.cuda.synchronize()
    out = [x.cpu().numpy() for x in m.images()]
    m.close()
    b = recon.BatchedOSEM(w, solves, cache_format=firlib.FIR_LEN_SNORM16,
                          cache_row_order=t.best["row_order"],
                          cache_lanes=(t.best["fwd_lanes"], 16), **kw)
    b.capture()
    b.iteration()
    torch.cuda.synchronize()
    ref = [x.cpu().numpy() for x in b.images()]
    b.close()
    assert all(np.isfinite(o).all() for o in out)
    diff = [i for i in range(len(out)) if not np.array_equal(out[i], ref[i])]
    assert not diff, diff


def test_multi_step_rejects_bad_jobs():
    """fir_osem_step_create refuses jobs sharing a workspace, and caches
    without the sheet form (and so without the slab form)."""
    w, solves = _c1_uneven_solves(2)
    ctx = firlib.context_for(w)
    ctx.set_cache_lanes(4, 16)
    s0 = solves[0]
    cache = ctx.path_cache(s0["vol"], s0["n_subsets"],
                           firlib.FIR_LEN_SNORM16,
                           row_order=firlib.CACHE_ROW_ORDER)
    nv = s0["vol"].n_voxels
    sens = torch.zeros(s0["n_subsets"] * nv, device="cuda")
    cache.sensitivity(sens)
    ws = torch.empty(cache.job_workspace() + 256, dtype=torch.uint8,
                     device="cuda")
    ws = ws[(-ws.data_ptr()) % 256:][:cache.job_workspace()]
    job = dict(cache=cache, subset=0, x=[torch.ones(nv, device="cuda")],
               y=[s0["y"]], bp=[torch.zeros(nv, device="cuda")],
               sens=sens[:nv], workspace=ws)
    st = firlib.OsemStep([job])
    st.run()
    st.close()
    with pytest.raises(firlib.FirError, match="share a workspace"):
        firlib.OsemStep([job, dict(job, subset=1)])
    ctx.set_cache_sheets(False)
    c2 = ctx.path_cache(s0["vol"], s0["n_subsets"], firlib.FIR_LEN_SNORM16,
                        row_order=firlib.CACHE_ROW_ORDER)
    assert not c2.forms()["sheets"]
    with pytest.raises(firlib.FirError, match="slab form"):
        firlib.OsemStep([dict(job, cache=c2)])
    c2.close()
    cache.close()
    ctx.close()
