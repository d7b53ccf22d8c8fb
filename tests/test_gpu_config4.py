"""Batched uneven reconstructions at full size: 64 OSEM solves of the
24-ring scanner's data (41.9 M LORs) in one HIP graph per iteration, on
mixed 128x128x47 and 160x160x47 grids with 8, 16 or 24 subsets (24 does not
divide the 224 views: interleaved uneven subsets, an extension), counts
log-uniform in [1e5, 1e8], 4 mm Gaussian every sub-iteration, path caches
in snorm16 lengths shared per grid.

Solve 0 carries the parity counts of the reference's 3 x 16 FIR_OSEM run
(tests/golden/c2_osem_golden.npz): after three batched iterations it must
be within 1e-4 relative L2 of the reference's image. One solve of every
(grid, subset count) group is checked against the same solve run alone on
the plain path (FP64 walk per LOR) for the same three iterations, and
against itself run alone on its group's cache (the per-image arithmetic of
the batch, to FP32 rounding).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_l2
from fir_amd import interfile as io
from fir_amd import lib as firlib
from fir_amd import recon
from fir_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N_SOLVES = 64
N_ITER = 3


def test_config4_uneven_batch_full_size():
    g = dict(np.load(os.path.join(GOLDEN, "c2_osem_golden.npz")))
    w = synth.c2()
    ctx = firlib.context_for(w)
    v160 = firlib.VolDesc.from_spec(w.vol)
    v128 = firlib.VolDesc.from_spec(io.VolSpec((128, 128, 47),
                                               (2.0, 2.0, 3.27),
                                               (-127.0, -127.0, 0.0)))
    act, hu = synth.phantom(w)
    y0 = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v160, torch.from_numpy(act.ravel()).cuda(), y0)
    mu = torch.from_numpy(hu.ravel()).cuda()
    ctx.hu_to_mu(mu)
    m = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v160, mu, m)
    counts, _ = synth.parity_counts(w, y0.cpu().numpy(), m.cpu().numpy())
    ctx.exp_clamped(m)
    y_ref = torch.from_numpy(counts).cuda()
    ctx.mul_clamped(y_ref, m)
    del m, mu, counts
    y0 = y0 / y0.sum()
    solves = [dict(vol=v160, n_subsets=16, y=y_ref)]
    for i in range(1, N_SOLVES):
        seed = 100 + i
        total = 10 ** np.random.default_rng(seed).uniform(5.0, 8.0)
        gen = torch.Generator(device="cuda").manual_seed(seed)
        y = torch.poisson(y0 * float(total), generator=gen)
        solves.append(dict(vol=v128 if i % 2 else v160,
                           n_subsets=(8, 16, 24)[i % 3], y=y))
    del y0
    kw = dict(fwhm=w.fwhm, conv_interval=1)
    b = recon.BatchedOSEM(w, solves, cache_format=firlib.FIR_LEN_SNORM16,
                          **kw)
    assert len(b.groups) == 6
    b.capture()          # iteration 1 eagerly, then captured
    for _ in range(N_ITER - 1):
        b.iteration()
    torch.cuda.synchronize()
    out = [x.cpu().numpy() for x in b.images()]
    assert all(np.isfinite(x).all() for x in out)

    e_ref = rel_l2(out[0], g["x_ref"])
    print(f"solve 0 (reference counts, 160 grid, 16 subsets): rel L2 "
          f"{e_ref:.2e} vs the reference", flush=True)
    assert e_ref < 1e-4

    for grp, members in zip(b.groups, b.members):
        i = members[-1]
        s = solves[i]
        alone = recon.OSEMGroup(b.ctx, s["vol"], s["n_subsets"], [s["y"]],
                                cache=grp.cache, **kw)
        for _ in range(N_ITER):
            alone.iteration_eager()
        c = firlib.context_for(w)
        c.set_uneven_subsets(True)
        plain = recon.OSEMGroup(c, s["vol"], s["n_subsets"], [s["y"]], **kw)
        for _ in range(N_ITER):
            plain.iteration_eager()
        torch.cuda.synchronize()
        e_alone = rel_l2(out[i], alone.x[0].cpu().numpy())
        e_plain = rel_l2(out[i], plain.x[0].cpu().numpy())
        print(f"solve {i} ({s['vol'].nx} grid, {s['n_subsets']} subsets, "
              f"{len(members)} in group): rel L2 {e_alone:.1e} vs alone on "
              f"the cache, {e_plain:.1e} vs the plain path", flush=True)
        assert e_alone < 1e-6 and e_plain < 1e-4, i
        c.close()
    b.close()
    ctx.close()


def test_config4_multi_step_equals_group_graphs():
    """The 64 solves advanced by fir_osem_step launches (every phase of
    every solve scheduled at a step in one launch, the projector phases
    persistent over one (job, family) work table) give every solve exactly
    the image the per-group graphs give it, over two iterations (the second
    replayed from the captured graph)."""
    w = synth.c2()
    ctx = firlib.context_for(w)
    v160 = firlib.VolDesc.from_spec(w.vol)
    v128 = firlib.VolDesc.from_spec(io.VolSpec((128, 128, 47),
                                               (2.0, 2.0, 3.27),
                                               (-127.0, -127.0, 0.0)))
    act, _ = synth.phantom(w)
    y0 = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v160, torch.from_numpy(act.ravel()).cuda(), y0)
    y0 = y0 / y0.sum()
    solves = []
    for i in range(N_SOLVES):
        seed = 100 + i
        total = 10 ** np.random.default_rng(seed).uniform(5.0, 8.0)
        gen = torch.Generator(device="cuda").manual_seed(seed)
        solves.append(dict(vol=v128 if i % 2 else v160,
                           n_subsets=(8, 16, 24)[i % 3],
                           y=torch.poisson(y0 * float(total), generator=gen)))
    del y0
    kw = dict(fwhm=w.fwhm, conv_interval=1)
    b = recon.BatchedOSEM(w, solves, cache_format=firlib.FIR_LEN_SNORM16,
                          **kw)
    b.capture()
    b.iteration()
    torch.cuda.synchronize()
    ref = [x.cpu().numpy() for x in b.images()]
    b.close()
    ctx.close()
    m = recon.MultiOSEM(w, solves, **kw)
    assert m.n_steps == 24
    m.capture()
    m.iteration()
    torch.cuda.synchronize()
    out = [x.cpu().numpy() for x in m.images()]
    m.close()
    diff = [i for i in range(N_SOLVES) if not np.array_equal(out[i], ref[i])]
    assert not diff, diff
