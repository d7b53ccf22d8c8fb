"""fir_amd.pyinteractor: the differentiable projector configured from the
header objects of the reference's PyInteractor (ScannerHeader, ProjHeader,
VolHeader: the same fields, scanner.py:47-62, proj.py:16-30, vol.py:22-31).
The headers here are plain objects with those fields (the reference's
package is not importable on the GPU host); the projector's results are
compared with what the reference's FIR_ForwardProj and FIR_BackProj wrote
for the same geometry (tests/golden/c1_golden.npz)."""
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import rel_max
from fir_amd import synth
from fir_amd.pyinteractor import proj_spec, scanner_spec, vol_spec


def c1_headers(offset=None):
    """The 4-ring x 128-crystal scanner, its span-1 sinogram template and
    the 64 x 64 x 8 image template, as the reference's PyInteractor holds
    them."""
    scanner = SimpleNamespace(
        name="c1", crystal_dimensions=[0.0, 0.0, 4.0],
        crystal_repeat_numbers=[1, 4], inter_crystal_distance=[0.0, 0.0],
        module_dimensions=[0.0, 0.0, 0.0], module_repeat_numbers=[1, 1],
        inter_module_distance=[0.0, 0.0], rsector_dimensions=[0.0, 0.0, 0.0],
        rsector_repeat_number=128, rsector_inner_radius=150.0)
    proj = SimpleNamespace(data_file_name=None, n_rings=4,
                           n_crystals_per_ring=128, segment_span=1,
                           n_segments=7, n_tang_coords=65)
    vol = SimpleNamespace(data_file_name=None, number_format="float",
                          n_bytes_per_pixel=4, vol_size=(64, 64, 8),
                          voxel_extent=(2.0, 2.0, 2.0), vol_offset=offset,
                          n_frames=1)
    return scanner, proj, vol


def test_specs_from_reference_headers():
    w = synth.c1()
    scanner, proj, vol = c1_headers()
    assert np.array_equal(scanner_spec(scanner).header_vector(),
                          w.scanner.header_vector())
    p = proj_spec(proj)
    assert (p.n_rings, p.n_crystals_per_ring, p.n_segments, p.n_tang_coords,
            p.segment_span) == (4, 128, 7, 65, 1)
    assert p.geometry() == w.proj.geometry()
    # an unset offset is VolHeader.check's centred default (vol.py:159-165)
    v = vol_spec(vol)
    assert tuple(v.size) == tuple(w.vol.size)
    assert tuple(v.offset) == (-63.0, -63.0, 0.0) == tuple(w.vol.offset)
    v2 = vol_spec(c1_headers(offset=(-10.0, -20.0, 5.0))[2])
    assert tuple(v2.offset) == (-10.0, -20.0, 5.0)


@pytest.mark.gpu
def test_torch_projector_matches_reference_executables(c1_golden):
    torch = pytest.importorskip("torch")
    from fir_amd.pyinteractor import TorchProjector
    proj = TorchProjector(*c1_headers())
    try:
        act = torch.from_numpy(c1_golden["act"].ravel()).cuda()
        # forward: what FIR_ForwardProj wrote, bit for bit
        ax = proj(act)
        assert np.array_equal(ax.cpu().numpy(), c1_golden["y0"].ravel())
        # adjoint: FIR_BackProj's image (its own atomic order is
        # non-deterministic at ~1e-7)
        y = torch.from_numpy(c1_golden["y"].ravel()).cuda()
        aty = proj.adjoint(y)
        assert rel_max(aty.cpu().numpy(), c1_golden["bp"].ravel()) < 1e-5
        # autograd: the vector-Jacobian product of the forward projector is
        # the backprojection kernel
        x = act.clone().requires_grad_()
        (g,) = torch.autograd.grad(proj(x), x, grad_outputs=y)
        assert rel_max(g.cpu().numpy(), aty.cpu().numpy()) < 1e-6
        # Poisson NLL gradient A^T(1 - y/(Ax + r))
        r = torch.full_like(ax, 0.05)
        xr = (act + 0.01).clone().requires_grad_()
        (gn,) = torch.autograd.grad(proj.nll(xr, y, r), xr)
        l = proj(act + 0.01) + r
        ref = proj.adjoint(torch.where(l > 0, 1.0 - y / l,
                                       torch.ones_like(l)))
        assert rel_max(gn.cpu().numpy(), ref.cpu().numpy()) < 1e-5
    finally:
        proj.close()
