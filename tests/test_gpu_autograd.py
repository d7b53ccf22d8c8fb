"""Differentiable projector (fir_amd.autograd): the backprojection kernel is
the adjoint of the forward kernel, the fused Poisson NLL gradient matches
central finite differences of the all-FP64 oracle and the backprojection
of the reference, and autograd composes with torch."""
import numpy as np
import pytest

import pyoracle as O
from conftest import rel_max
from fir_amd import autograd as fad
from fir_amd import lib as firlib
from fir_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def c2proj():
    return fad.Projector(synth.c2())


def test_adjoint_dot_product_batch(c2proj):
    """<A x, y> = <x, A^T y> for a batch of 4 random images, through
    torch.autograd (the vector-Jacobian product is A^T y)."""
    g = torch.Generator(device="cuda").manual_seed(2)
    nv, nb = c2proj.vol.n_voxels, c2proj.ctx.n_bins
    x = torch.rand(4, nv, device="cuda", generator=g).requires_grad_()
    y = torch.rand(4, nb, device="cuda", generator=g)
    ax = c2proj(x)
    (aty,) = torch.autograd.grad(ax, x, grad_outputs=y)
    for b in range(4):
        lhs = float((ax[b].double() * y[b].double()).sum())
        rhs = float((x[b].detach().double() * aty[b].double()).sum())
        assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_nll_gradient_matches_backprojection(c2proj):
    """grad = A^T(1 - y/(Ax + r)) equals the backprojection kernel applied
    to the weights computed from the forward kernel."""
    w = synth.c2()
    act, _ = synth.phantom(w)
    x = torch.from_numpy(act.ravel()).cuda() + 0.05
    ax = c2proj(x)
    g = torch.Generator(device="cuda").manual_seed(3)
    y = torch.poisson(ax * 2.0, generator=g)
    r = torch.full_like(ax, 0.1)
    xr = x.clone().requires_grad_()
    loss = c2proj.nll(xr, y, r)
    (grad,) = torch.autograd.grad(loss, xr)
    l = ax + r
    wts = torch.where(l > 0, 1.0 - y / l, torch.ones_like(l))
    ref = torch.zeros_like(x)
    c2proj.ctx.backward_subset(c2proj.vol, wts.contiguous(), ref, 0, 1)
    torch.cuda.synchronize()
    assert rel_max(grad.cpu().numpy(), ref.cpu().numpy()) < 1e-5
    lref = float((l.double() - y.double() * torch.log(l.double())).sum())
    assert abs(float(loss) - lref) <= 1e-9 * abs(lref)


def test_nll_gradient_finite_differences():
    """Central differences (eps = 1e-3) of the all-FP64 oracle NLL at 64
    random voxels of the 4x128 workload."""
    w = synth.c1()
    P = fad.Projector(w)
    xy, z = O.scanner_tables(w.scanner)
    act, _ = synth.phantom(w)
    x = act.ravel().astype(np.float64) + 0.2
    y0 = O.forward(w.vol, w.proj, xy, z, x)
    y = np.random.default_rng(4).poisson(y0 * 3.0).astype(np.float64)
    r = np.full_like(y, 0.05)

    def nll64(xv):
        l = O.forward(w.vol, w.proj, xy, z, xv) + r
        return float(np.sum(l - y * np.log(l)))

    xt = torch.from_numpy(x.astype(np.float32)).cuda().requires_grad_()
    loss = P.nll(xt, torch.from_numpy(y.astype(np.float32)).cuda(),
                 torch.from_numpy(r.astype(np.float32)).cuda())
    (grad,) = torch.autograd.grad(loss, xt)
    grad = grad.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(5)
    sens = O.backward(w.vol, w.proj, xy, z, np.ones_like(y))
    cand = np.nonzero(sens > 0.1 * sens.max())[0]
    eps = 1e-3
    for v in rng.choice(cand, 64, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[v] += eps
        xm[v] -= eps
        fd = (nll64(xp) - nll64(xm)) / (2 * eps)
        assert abs(grad[v] - fd) <= 2e-3 * max(1.0, abs(fd)), (v, grad[v], fd)


def test_gradient_descent_step_decreases_nll(c2proj):
    w = synth.c2()
    act, _ = synth.phantom(w)
    truth = torch.from_numpy(act.ravel()).cuda()
    y = c2proj(truth).detach()
    x = torch.ones_like(truth).requires_grad_()
    l0 = c2proj.nll(x, y)
    (g,) = torch.autograd.grad(l0, x)
    x1 = (x - 1e-4 * g).clamp_min(0).detach()
    assert float(c2proj.nll(x1, y)) < float(l0)
