"""fir_convolve_noalloc (the Gaussian through the library's static device
scratch, no device allocation) against fir_convolve, through the C-ABI:
the same kernels and taps, so the frames must be identical bit for bit.
Two calls in a row on one context, the second on a smaller volume (the
static buffer keeps the first call's contents), then a volume beyond the
buffer's limit, which must fail before anything is launched."""
import numpy as np
import pytest

from fir_amd import interfile as io
from fir_amd import lib as firlib

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# 24-ring image (two frames, 4 mm FWHM, 150 mm edge restore), then the
# 4-ring one (one frame, 4 mm, no restore) and a 2.5 mm grid with 6 mm FWHM
CASES = [
    (io.VolSpec((160, 160, 47), (2.0, 2.0, 3.27), (-159.0, -159.0, 0.0)),
     2, (4.0, 4.0, 4.0), 150.0),
    (io.VolSpec((64, 64, 8), (2.0, 2.0, 2.0), (-63.0, -63.0, 0.0)),
     1, (4.0, 4.0, 4.0), 0.0),
    (io.VolSpec((48, 48, 8), (2.5, 2.5, 2.0), (-58.75, -58.75, 0.0)),
     3, (6.0, 6.0, 5.0), 50.0),
]


def _images(spec, frames, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    n = spec.n_voxels * frames
    return torch.rand(n, generator=g, device="cuda") * 10.0


def test_noalloc_gaussian_equals_convolve():
    ctx = firlib.Context(0)
    ctx.set_stream(torch.cuda.current_stream())
    for k, (spec, frames, fwhm, cut) in enumerate(CASES):
        v = firlib.VolDesc.from_spec(spec)
        img = _images(spec, frames, 100 + k)
        a = img.clone()
        b = img.clone()
        print(f"case {k}: {spec.size} x {frames} frames, fwhm {fwhm}, "
              f"cut {cut}", flush=True)
        ctx.convolve_noalloc(v, a, fwhm, cut, n_frames=frames)
        ctx.convolve(v, b, fwhm, cut, n_frames=frames)
        torch.cuda.synchronize()
        assert not torch.equal(a, img), k      # the Gaussian ran
        assert torch.equal(a, b), k
    ctx.close()


def test_noalloc_gaussian_first_call_on_fresh_context():
    """The static path on a context that has never run fir_convolve (no
    context scratch or taps exist), then fir_convolve on a second context
    for the check."""
    spec, frames, fwhm, cut = CASES[1]
    v = firlib.VolDesc.from_spec(spec)
    img = _images(spec, frames, 7)
    a, b = img.clone(), img.clone()
    c1 = firlib.Context(0)
    c1.set_stream(torch.cuda.current_stream())
    c1.convolve_noalloc(v, a, fwhm, cut, n_frames=frames)
    c2 = firlib.Context(0)
    c2.set_stream(torch.cuda.current_stream())
    c2.convolve(v, b, fwhm, cut, n_frames=frames)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    c1.close()
    c2.close()


def test_noalloc_gaussian_rejects_oversized_frames():
    ctx = firlib.Context(0)
    v = firlib.VolDesc.from_spec(io.VolSpec((512, 512, 64), (1.0, 1.0, 1.0),
                                            (-255.5, -255.5, 0.0)))
    img = torch.zeros(16, device="cuda")
    with pytest.raises(firlib.FirError, match="static scratch"):
        ctx.convolve_noalloc(v, img, (4.0, 4.0, 4.0), 0.0)
    ctx.close()
