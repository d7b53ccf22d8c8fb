"""Differentiable projector over the C-ABI.

* ``ForwardProjection`` -- torch.autograd.Function: forward = the forward
  projector (fir_forward), backward = its exact adjoint, the
  backprojection kernel (fir_backward_subset with one subset).
* ``PoissonNLL`` -- torch.autograd.Function for
  L(x) = sum_b ((Ax + r)_b - y_b log (Ax + r)_b) whose gradient
  A^T (1 - y / (Ax + r)) comes from the fused kernel fir_poisson_nll_grad.
* ``Projector`` -- holds the fir context and volume descriptor and applies
  both to single images or batches (B, N_vox).

The reference's OSEM needs no gradient; this extends its projector pair
(projections.cc:15-184) to gradient-based reconstruction.
"""
from __future__ import annotations

import torch

from . import lib as firlib


def _as_batch(t):
    return (t.unsqueeze(0), True) if t.dim() == 1 else (t, False)


class ForwardProjection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, fir_ctx, vol):
        x, single = _as_batch(img.contiguous().float())
        out = torch.empty(x.shape[0], fir_ctx.n_bins, device=x.device,
                          dtype=torch.float32)
        for b in range(x.shape[0]):
            fir_ctx.forward(vol, x[b], out[b])
        ctx.fir_ctx, ctx.vol, ctx.single = fir_ctx, vol, single
        return out[0] if single else out

    @staticmethod
    def backward(ctx, grad_out):
        g, _ = _as_batch(grad_out.contiguous().float())
        res = torch.zeros(g.shape[0], ctx.vol.n_voxels, device=g.device,
                          dtype=torch.float32)
        for b in range(g.shape[0]):
            ctx.fir_ctx.backward_subset(ctx.vol, g[b], res[b], 0, 1)
        return (res[0] if ctx.single else res), None, None


class PoissonNLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, y, r, fir_ctx, vol):
        x, single = _as_batch(img.contiguous().float())
        yb, _ = _as_batch(y.contiguous().float())
        rb = None
        if r is not None:
            rb, _ = _as_batch(r.contiguous().float())
        n = x.shape[0]
        grads = torch.empty(n, vol.n_voxels, device=x.device,
                            dtype=torch.float32)
        losses = torch.empty(n, device=x.device, dtype=torch.float64)
        for b in range(n):
            fir_ctx.poisson_nll_grad(
                vol, x[b], yb[b if yb.shape[0] > 1 else 0],
                None if rb is None else rb[b if rb.shape[0] > 1 else 0],
                grads[b], losses[b:b + 1])
        ctx.save_for_backward(grads)
        ctx.single = single
        return losses[0] if single else losses

    @staticmethod
    def backward(ctx, grad_loss):
        (grads,) = ctx.saved_tensors
        g = grad_loss.reshape(-1, 1).to(grads.dtype) * grads
        return (g[0] if ctx.single else g), None, None, None, None


class Projector:
    """Forward projector A and Poisson NLL for one geometry."""

    def __init__(self, workload=None, fir_ctx=None, vol=None, device=0,
                 geometric=False):
        if workload is not None:
            self.ctx = firlib.context_for(workload, device, geometric)
            self.vol = firlib.VolDesc.from_spec(workload.vol)
        else:
            self.ctx, self.vol = fir_ctx, vol

    def __call__(self, img):
        return ForwardProjection.apply(img, self.ctx, self.vol)

    def nll(self, img, y, r=None):
        return PoissonNLL.apply(img, y, r, self.ctx, self.vol)
