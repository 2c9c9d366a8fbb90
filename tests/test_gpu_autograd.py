"""GPU parity of the autograd bindings (ola, tv, OlaLoss) against the FP64
loss oracle: values to 1e-5, gradients to 1e-4 (relative L2), on the 96^2,
3x3-patch, 9x9-PSF geometry."""
import numpy as np
import pytest

from oracle import loss_oracle as L
from oracle import ola_oracle as O
from tests.helpers import case_inputs, rel_l2

pytestmark = pytest.mark.gpu

SX, CSF, SF = (96, 96), (3, 3), (9, 9)


def _t(a, cuda, grad=False):
    import torch
    t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=cuda)
    return t.requires_grad_(grad)


def _n(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


@pytest.fixture(scope='module')
def setup(cuda):
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    wa = win2winaux(SX, CSF, 0.5)
    plan = OlaPlan(SX, SF, wa, 'valid', max_frames=2)
    x, f, g = case_inputs(31, SX, CSF, SF)
    owa = O.Win2WinAux(SX, CSF, 0.5)
    y = O.forward(x, f, owa) + 0.1 * g
    return plan, owa, _f32(x), _f32(f), _f32(y)


def test_tv_value_and_grad(cuda):
    import torch
    from vmbd_amd.autograd import tv
    print('case tv', flush=True)
    x = _f32(np.random.default_rng(3).random((2, 40, 56)))
    xt = _t(x, cuda, True)
    v = tv(xt, 0.05)
    v.backward(torch.tensor(2.5, device=cuda))
    assert abs(v.item() - L.tv(x, 0.05)) < 1e-5 * L.tv(x, 0.05)
    assert rel_l2(_n(xt.grad), 2.5 * L.tv_grad(x, 0.05)) < 1e-5


def test_loss_grads_vs_oracle(cuda, setup):
    from vmbd_amd.autograd import OlaLoss
    print('case OlaLoss grads', flush=True)
    plan, owa, x, f, y = setup
    lam, eps = 1e-3, 1e-2
    xt, ft = _t(x, cuda, True), _t(f, cuda, True)
    loss = OlaLoss(plan, lam, eps)(xt, ft, _t(y, cuda))
    loss.backward()
    ref = L.loss(x, f, y, owa, lam, eps)
    gx, gf = L.loss_grad(x, f, y, owa, lam, eps)
    assert abs(loss.item() - ref) < 1e-5 * ref
    assert rel_l2(_n(xt.grad), gx) < 1e-4
    assert rel_l2(_n(ft.grad), gf) < 1e-4
    # directional derivative of the FP64 loss (central differences) vs the device gradient
    rng = np.random.default_rng(4)
    vx, vf = rng.standard_normal(x.shape), rng.standard_normal(f.shape)
    h = 3e-6
    fd = (L.loss(x + h * vx, f + h * vf, y, owa, lam, eps) - L.loss(x - h * vx, f - h * vf, y, owa, lam, eps)) / (2 * h)
    an = np.sum(_n(xt.grad) * vx) + np.sum(_n(ft.grad) * vf)
    assert abs(an - fd) < 1e-4 * abs(fd)


def test_single_input_grads_match_joint(cuda, setup):
    import torch
    from vmbd_amd.autograd import ola
    print('case ola single-input grads', flush=True)
    plan, owa, x, f, y = setup
    yt = _t(y, cuda)
    xa, fa = _t(x, cuda, True), _t(f, cuda, True)
    (0.5 * ((ola(xa, fa, plan) - yt) ** 2).sum()).backward()
    xb = _t(x, cuda, True)
    (0.5 * ((ola(xb, _t(f, cuda), plan) - yt) ** 2).sum()).backward()
    fb = _t(f, cuda, True)
    (0.5 * ((ola(_t(x, cuda), fb, plan) - yt) ** 2).sum()).backward()
    assert torch.equal(xa.grad, xb.grad)
    assert torch.equal(fa.grad, fb.grad)


def test_batched_frames(cuda, setup):
    from vmbd_amd.autograd import OlaLoss
    print('case OlaLoss batched', flush=True)
    plan, owa, x, f, y = setup
    x2 = np.stack([x, x[::-1].copy()])
    f2 = np.stack([f, f[::-1].copy()])
    y2 = np.stack([y, y[:, ::-1].copy()])
    xt, ft = _t(x2, cuda, True), _t(f2, cuda, True)
    OlaLoss(plan, 1e-3, 1e-2)(xt, ft, _t(y2, cuda)).backward()
    for b in range(2):
        gx, gf = L.loss_grad(x2[b], f2[b], y2[b], owa, 1e-3, 1e-2)
        assert rel_l2(_n(xt.grad[b]), gx) < 1e-4
        assert rel_l2(_n(ft.grad[b]), gf) < 1e-4


def test_adam_decreases_loss(cuda, setup):
    import torch
    from vmbd_amd.autograd import OlaLoss
    print('case adam', flush=True)
    plan, owa, x, f, y = setup
    xt = _t(np.clip(x + 0.1, 0, None), cuda, True)
    ft = _t(f, cuda, True)
    opt = torch.optim.Adam([xt, ft], lr=1e-3)
    crit = OlaLoss(plan, 1e-3, 1e-2)
    yt = _t(y, cuda)
    first = None
    for _ in range(20):
        opt.zero_grad()
        loss = crit(xt, ft, yt)
        loss.backward()
        opt.step()
        first = loss.item() if first is None else first
    assert crit(xt, ft, yt).item() < 0.9 * first
