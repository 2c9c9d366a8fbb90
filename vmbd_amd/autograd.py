"""torch.autograd bindings of the EFF/OLA operator and of the TV prior.

    plan = OlaPlan(sx, sf, winaux)
    y    = ola(x, f, plan)                 # B(x, f), differentiable in x and f
    loss = 0.5 * ((y - obs) ** 2).sum() + lam * tv(x, eps)
    loss.backward()

The backward of ``ola`` is the exact adjoint pair (x-adjoint F'g and
f-adjoint X'g, olaGPU.py:261-311 without the patch-local truncation), run as
one fused launch sequence when both gradients are needed; it reuses the row
transforms of the forward when nothing else ran on the plan in between.
``tv`` is isotropic and eps-smoothed (the reference has no TV prior):
TV(x) = sum sqrt(dx^2 + dy^2 + eps^2), forward differences, zero past the
last row/column; its value and gradient are HIP kernels, and the gradient
reads the upstream scalar from device memory (no host synchronisation).
"""
import torch

from . import _lib
from .optGPU import _Scratch


class _OlaFunction(torch.autograd.Function):

    @staticmethod
    def forward(ctx, x, f, plan, exact):
        y = plan.forward(x, f)
        ctx.plan = plan
        ctx.exact = bool(exact)
        ctx.save_for_backward(x, f)
        return y

    @staticmethod
    def backward(ctx, gy):
        x, f = ctx.saved_tensors
        plan = ctx.plan
        gy = gy.contiguous()
        need_x, need_f = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        gx = gf = None
        if need_x and need_f:
            gx, gf = plan.adjoint_xf(gy, x, f, exact=ctx.exact)
        elif need_x:
            gx = plan.adjoint_x(gy, f, exact=ctx.exact)
        elif need_f:
            gf = plan.adjoint_f(gy, x, exact=ctx.exact)
        return gx, gf, None, None


def ola(x, f, plan, exact=True):
    """Differentiable B(x, f) on `plan` (x: ([B,] H, W), f: ([B,] P, sf0, sf1))."""
    return _OlaFunction.apply(x, f, plan, exact)


class _TvFunction(torch.autograd.Function):

    @staticmethod
    def forward(ctx, x, eps):
        if x.dtype != torch.float32 or x.device.type != 'cuda':
            raise TypeError('tv: expected a float32 device tensor')
        xc = x.contiguous()
        H, W = int(xc.shape[-2]), int(xc.shape[-1])
        K = xc.numel() // (H * W)
        sc = _Scratch(xc.numel(), xc.device)
        val = torch.empty(1, dtype=torch.float64, device=xc.device)
        _lib.call('vmbd_tv', _lib.ptr(xc), K, H, W, float(eps), _lib.ptr(sc.partials), _lib.ptr(val),
                  _lib.stream_ptr())
        ctx.eps = float(eps)
        ctx.save_for_backward(xc)
        return val.to(torch.float32).reshape(())

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        H, W = int(x.shape[-2]), int(x.shape[-1])
        K = x.numel() // (H * W)
        scale = g.to(torch.float32).contiguous().reshape(1)
        out = torch.empty_like(x)
        _lib.call('vmbd_tv_grad', _lib.ptr(out), _lib.ptr(x), K, H, W, ctx.eps, _lib.ptr(scale), 1.0, 0,
                  _lib.stream_ptr())
        return out, None


def tv(x, eps=1e-3):
    """Isotropic eps-smoothed total variation, summed over all images of x."""
    return _TvFunction.apply(x, eps)


class OlaLoss(torch.nn.Module):
    """L(x, f) = 1/2 |B(x, f) - y|^2 + lam TV(x)."""

    def __init__(self, plan, lam=1e-3, eps=1e-3, exact=True):
        super().__init__()
        self.plan = plan
        self.lam = float(lam)
        self.eps = float(eps)
        self.exact = bool(exact)

    def forward(self, x, f, y):
        r = ola(x, f, self.plan, self.exact) - y
        out = 0.5 * (r * r).sum()
        if self.lam:
            out = out + self.lam * tv(x, self.eps)
        return out
