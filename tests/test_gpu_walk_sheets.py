"""The walked OSEM projector family by family (fir_set_walk_sheets): one
block per (view, tangential bin) stages that line's sheet of the image in
LDS for the line integrals and accumulates the family's backprojection
there. The line integrals use the same values in the same order as the
LOR-by-LOR walk, so the ratios are identical; the backprojection differs
only by the order of its FP32 additions (atomics in both forms): bar 1e-6
relative. Results also match the reference's OSEM (tests/golden)."""
import numpy as np
import pytest

from conftest import rel_l2
from fir_amd import lib as firlib
from fir_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def project_both(w, x, y, bias, subsets, S, geometric=False):
    ctx = firlib.context_for(w, geometric=geometric)
    v = firlib.VolDesc.from_spec(w.vol)
    out = []
    for on in (True, False):
        ctx.set_walk_sheets(on)
        res = []
        for s in subsets:
            bp = torch.zeros(v.n_voxels, device="cuda")
            ctx.osem_project(v, x, y, bias, bp, s, S)
            res.append(bp.cpu().numpy())
        out.append(res)
    ctx.close()
    return out


def test_family_projector_matches_lor_walk_c1(c1_golden):
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    w = synth.c1()
    x = torch.from_numpy(c1_golden["act"].ravel() + 0.1).cuda()
    y = torch.from_numpy(c1_golden["y"]).cuda()
    bias = torch.full_like(y, 0.25)
    for b in (None, bias):
        fam, lor = project_both(w, x, y, b, range(4), 4)
        for a, c in zip(fam, lor):
            assert np.abs(a).max() > 0
            assert rel_l2(a, c) < 1e-6


def test_family_projector_matches_lor_walk_c2():
    """Full scanner (24 rings x 448 crystals, 160 x 160 x 47): two subsets,
    reference and geometric walks."""
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    w = synth.c2()
    act, _ = synth.phantom(w)
    x = torch.from_numpy(act.ravel()).cuda() + 0.1
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    y = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v, x, y)
    ctx.close()
    y = torch.poisson(y * 5.0, generator=torch.Generator(device="cuda")
                      .manual_seed(3))
    for geometric in (False, True):
        fam, lor = project_both(w, x, y, None, (0, 9), 16, geometric)
        for a, c in zip(fam, lor):
            assert rel_l2(a, c) < 1e-6


@pytest.mark.parametrize("with_sino", [True, False])
def test_family_backprojection_matches_lor_walk(c1_golden, with_sino):
    """fir_backward / sensitivity family by family against the LOR-by-LOR
    scatter, C1 (all subsets) and C2 (one pass of every LOR)."""
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    for w, y in ((synth.c1(), c1_golden["y"]), (synth.c2(), None)):
        ctx = firlib.context_for(w)
        v = firlib.VolDesc.from_spec(w.vol)
        if y is None:
            y = np.random.default_rng(4).poisson(
                3.0, ctx.n_bins).astype(np.float32)
        yt = torch.from_numpy(y).cuda() if with_sino else None
        out = []
        for on in (True, False):
            ctx.set_walk_sheets(on)
            f = torch.zeros(4 * v.n_voxels, device="cuda")
            if with_sino:
                ctx.backward(v, yt, f, 4)
            else:
                ctx.sensitivity(v, f, 4)
            out.append(f.cpu().numpy())
        assert np.abs(out[0]).max() > 0
        assert rel_l2(out[0], out[1]) < 1e-6
        ctx.close()


def test_forward_unaffected_by_walk_sheets(c1_golden):
    """fir_forward walks LOR by LOR whatever the setting (a forward pass is
    too light to repay a family's staging): bit for bit the reference's
    forward projection at C1, and the same output either way at C2, every
    bin written."""
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible")
    w = synth.c1()
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    x = torch.from_numpy(c1_golden["act"].ravel()).cuda()
    s = torch.zeros(ctx.n_bins, device="cuda")
    ctx.forward(v, x, s)
    assert np.array_equal(s.cpu().numpy(), c1_golden["y0"])
    ctx.close()
    w = synth.c2()
    ctx = firlib.context_for(w)
    v = firlib.VolDesc.from_spec(w.vol)
    act, _ = synth.phantom(w)
    x = torch.from_numpy(act.ravel()).cuda()
    out = []
    for on in (True, False):
        ctx.set_walk_sheets(on)
        s = torch.full((ctx.n_bins,), -1.0, device="cuda")
        ctx.forward(v, x, s)
        out.append(s.cpu().numpy())
    assert np.array_equal(out[0], out[1]) and (out[0] > 0).any()
    ctx.close()
