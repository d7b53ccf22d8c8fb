"""torch.autograd projector pair for callers of the reference's PyInteractor.

The reference's PyInteractor projects by writing Interfile files and
running the executables: ``forward`` (PyInteractor/forward.py:19-72) runs
FIR_ForwardProj, ``backward`` (PyInteractor/backward.py:16-64) runs
FIR_BackProj. ``TorchProjector`` takes the same header objects -- a
ScannerHeader (scanner.py:47-62), a projection template ProjHeader
(proj.py:16-30) and a volume template VolHeader (vol.py:22-31), or any
objects with those fields -- and applies the same projector pair to device
tensors in memory through the C-ABI:

* ``proj(img)``: image (N_vox,) or (B, N_vox), x fastest as in the volume's
  .i33 file, to the sinogram (n_bins,) or (B, n_bins) in the .s file's
  segment order (what FIR_ForwardProj writes); differentiable, its
  gradient is the exact adjoint, the backprojection kernel;
* ``proj.adjoint(sino)``: the backprojection (FIR_BackProj's result);
* ``proj.nll(img, y, r)``: the Poisson log-likelihood
  sum((Ax + r) - y log(Ax + r)), whose gradient A^T(1 - y/(Ax + r)) comes
  from the fused gradient kernel.

The reference itself offers no gradient; OSEM stays on the executables'
plain path (FIR_OSEM).
"""
from __future__ import annotations

from types import SimpleNamespace

from .. import interfile
from .. import lib as firlib
from ..autograd import ForwardProjection, PoissonNLL


def scanner_spec(scanner) -> interfile.ScannerSpec:
    """The scanner keys of a ScannerHeader (scanner.py:47-62)."""
    return interfile.ScannerSpec(
        rsector_repeat_number=int(scanner.rsector_repeat_number),
        rsector_inner_radius=float(scanner.rsector_inner_radius),
        crystal_dimensions=tuple(float(v) for v in scanner.crystal_dimensions),
        crystal_repeat_numbers=tuple(
            int(v) for v in scanner.crystal_repeat_numbers),
        inter_crystal_distance=tuple(
            float(v) for v in scanner.inter_crystal_distance),
        module_dimensions=tuple(float(v) for v in scanner.module_dimensions),
        module_repeat_numbers=tuple(
            int(v) for v in scanner.module_repeat_numbers),
        inter_module_distance=tuple(
            float(v) for v in scanner.inter_module_distance),
        rsector_dimensions=tuple(float(v) for v in scanner.rsector_dimensions))


def proj_spec(proj) -> interfile.ProjSpec:
    """The sinogram keys of a ProjHeader template (proj.py:16-30)."""
    return interfile.ProjSpec(
        n_rings=int(proj.n_rings),
        n_crystals_per_ring=int(proj.n_crystals_per_ring),
        n_segments=int(proj.n_segments),
        n_tang_coords=int(proj.n_tang_coords),
        segment_span=int(proj.segment_span))


def vol_spec(vol) -> interfile.VolSpec:
    """The image geometry of a VolHeader template (vol.py:22-31); an unset
    offset is the centred default of VolHeader.check (vol.py:159-165)."""
    size = tuple(int(v) for v in vol.vol_size)
    voxel = tuple(float(v) for v in vol.voxel_extent)
    offset = vol.vol_offset
    if offset is None:
        offset = (-(size[0] - 1) * voxel[0] / 2,
                  -(size[1] - 1) * voxel[1] / 2, 0.0)
    return interfile.VolSpec(size=size, voxel=voxel,
                             offset=tuple(float(v) for v in offset))


class TorchProjector:
    """Forward projector A of one (scanner, sinogram, image) geometry on a
    GPU, with its adjoint and the Poisson log-likelihood."""

    def __init__(self, scanner, proj_template, vol_template, device: int = 0):
        geom = SimpleNamespace(scanner=scanner_spec(scanner),
                               proj=proj_spec(proj_template))
        self.vol_spec = vol_spec(vol_template)
        self.ctx = firlib.context_for(geom, device)
        self.vol = firlib.VolDesc.from_spec(self.vol_spec)
        self.n_bins = self.ctx.n_bins
        self.n_voxels = self.vol.n_voxels

    def __call__(self, img):
        return ForwardProjection.apply(img, self.ctx, self.vol)

    def adjoint(self, sino):
        """A^T sino (the backprojection), (n_bins,) or (B, n_bins)."""
        import torch
        g = sino.contiguous().float()
        single = g.dim() == 1
        g2 = g.unsqueeze(0) if single else g
        out = torch.zeros(g2.shape[0], self.n_voxels, device=g.device,
                          dtype=torch.float32)
        for b in range(g2.shape[0]):
            self.ctx.backward_subset(self.vol, g2[b], out[b], 0, 1)
        return out[0] if single else out

    def nll(self, img, y, r=None):
        return PoissonNLL.apply(img, y, r, self.ctx, self.vol)

    def close(self):
        self.ctx.close()
