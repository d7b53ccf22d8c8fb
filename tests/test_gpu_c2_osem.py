"""The 24-ring workload's full OSEM schedule against the reference.

Data: the seeded 12-ellipsoid phantom of the 24 rings x 448 crystals
scanner (41 932 800 LORs, 160x160x47 image at 2x2x3.27 mm), forward
projected, attenuated by its 200 mm water cylinder and drawn as Poisson
counts at 5e8 (synth.parity_counts: bit-reproducible on any host, given
bit-identical forward projections). The golden image is the reference's
FIR_OSEM on the same counts (3 iterations x 16 subsets, ACF from the HU
map, 4 mm FWHM Gaussian every sub-iteration, no cut; OSEM.cc:98-383,
reconAlgos.cc:63-201), made by tests/golden/make_golden.py c2osem, beside
the same run through the all-FP64 oracle.

Checks:
* the full forward projections of the phantom and of its mu map are
  bit-identical to the reference's (SHA-256 of all 41.9 M bins), and so are
  the counts drawn from them;
* the plain path (FP64 walk per LOR) and the path cache in both 16-bit
  length formats (the bench's path: sheet-staged line integrals and the
  family backprojection, sensitivity from the same cache) land within
  1e-4 relative L2 of the reference's image and of the FP64 oracle's;
* a batch of 8 images on a 4-lane cache (line integrals from the sheets
  of all 8 images staged in LDS slab by slab, 576 rows and about a dozen
  slabs per family) gives every image exactly its one-image pass, and so
  does a batch of 4.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_l2
from fir_amd import lib as firlib
from fir_amd import recon
from fir_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

BAR = 1e-4


def _sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def c2():
    g = dict(np.load(os.path.join(GOLDEN, "c2_osem_golden.npz")))
    w = synth.c2()
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    act, hu = synth.phantom(w)
    y0 = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v, torch.from_numpy(act.ravel()).cuda(), y0)
    mu = torch.from_numpy(hu.ravel()).cuda()
    ctx.hu_to_mu(mu)
    m = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v, mu, m)
    y0h, mh = y0.cpu().numpy(), m.cpu().numpy()
    counts, scale = synth.parity_counts(w, y0h, mh)
    acf = m.clone()
    ctx.exp_clamped(acf)
    y = torch.from_numpy(counts).cuda()
    ctx.mul_clamped(y, acf)   # FIR_OSEM: inputProj *= ACF (OSEM.cc:336-337)
    torch.cuda.synchronize()
    d = dict(g=g, w=w, ctx=ctx, v=v, y=y, hashes=(_sha(y0h), _sha(mh),
                                                   _sha(counts)),
             scale=scale)
    del y0, m, acf, mu
    yield d
    ctx.close()


def _run(ctx, v, w, y, cache=None):
    grp = recon.OSEMGroup(ctx, v, w.n_subsets, [y], fwhm=w.fwhm,
                          conv_interval=1, cache=cache)
    for _ in range(w.n_iterations):
        grp.iteration_eager()
    torch.cuda.synchronize()
    return grp.x[0].cpu().numpy()


def test_c2_forward_projections_and_counts_bit_exact(c2):
    g = c2["g"]
    y0_sha, m_sha, c_sha = c2["hashes"]
    assert y0_sha == str(g["y0_sha256"]), "phantom forward projection"
    assert m_sha == str(g["m_sha256"]), "mu-map forward projection"
    assert c2["scale"] == float(g["scale"])
    assert c_sha == str(g["counts_sha256"]), "counts"


def test_c2_osem_plain_vs_reference(c2):
    x = _run(c2["ctx"], c2["v"], c2["w"], c2["y"])
    e_ref = rel_l2(x, c2["g"]["x_ref"])
    e_64 = rel_l2(x, c2["g"]["x_f64"])
    print(f"plain path: rel L2 {e_ref:.3e} vs reference, {e_64:.3e} vs "
          f"FP64 oracle", flush=True)
    assert e_ref < BAR and e_64 < BAR


@pytest.mark.parametrize("fmt", [firlib.FIR_LEN_SNORM16,
                                 firlib.FIR_LEN_FP16])
def test_c2_osem_path_cache_vs_reference(c2, fmt):
    ctx, v, w = c2["ctx"], c2["v"], c2["w"]
    cache = ctx.path_cache(v, w.n_subsets, fmt,
                           row_order=firlib.CACHE_ROW_ORDER)
    forms = cache.forms()
    assert forms["sheets"] and forms["family_lists"], forms
    x = _run(ctx, v, w, c2["y"], cache=cache)
    cache.close()
    e_ref = rel_l2(x, c2["g"]["x_ref"])
    e_64 = rel_l2(x, c2["g"]["x_f64"])
    print(f"path cache fmt {fmt}: rel L2 {e_ref:.3e} vs reference, "
          f"{e_64:.3e} vs FP64 oracle", flush=True)
    assert e_ref < BAR and e_64 < BAR


def test_c2_slab_batches_equal_single_image_passes(c2):
    ctx, v, w = c2["ctx"], c2["v"], c2["w"]
    ctx.set_cache_lanes(4, 16)
    try:
        cache = ctx.path_cache(v, w.n_subsets, firlib.FIR_LEN_SNORM16,
                               row_order=firlib.CACHE_ROW_ORDER)
    finally:
        ctx.set_cache_lanes(8, 16)
    assert cache.forms()["slabs"]
    rng = np.random.default_rng(12)
    n = v.n_voxels
    xs = [torch.from_numpy(rng.uniform(0.5, 2.0, n).astype(np.float32))
          .cuda() for _ in range(8)]
    ys = [c2["y"] * float(s) for s in rng.uniform(0.05, 2.0, 8)]
    for nb in (8, 4):
        bps = [torch.zeros(n, device="cuda") for _ in range(nb)]
        cache.osem_project(xs[:nb], ys[:nb], None, bps, 5)
        for x, y, bp in zip(xs[:nb], ys[:nb], bps):
            one = torch.zeros(n, device="cuda")
            cache.osem_project([x], [y], None, [one], 5)
            torch.cuda.synchronize()
            assert torch.equal(bp, one), nb
    cache.close()
