"""Pins of the FP64 loss oracle: gradients against central finite
differences, TV against its closed form on simple images."""
import numpy as np

from oracle import loss_oracle as L
from oracle import ola_oracle as O
from tests.helpers import case_inputs


def test_tv_closed_forms():
    eps = 1e-3
    assert abs(L.tv(np.full((5, 6), 0.3), eps) - 30 * eps) < 1e-15          # constant: eps per pixel
    ramp = np.tile(np.arange(6.0), (5, 1))                                   # unit slope along columns
    ref = 5 * 5 * np.sqrt(1 + eps * eps) + 5 * eps                           # last column: zero difference
    assert abs(L.tv(ramp, eps) - ref) < 1e-12


def test_tv_grad_finite_differences():
    rng = np.random.default_rng(0)
    x = rng.random((2, 7, 9))
    eps = 0.05
    g = L.tv_grad(x, eps)
    h = 1e-6
    for _ in range(20):
        idx = tuple(rng.integers(0, s) for s in x.shape)
        xp = x.copy()
        xm = x.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = (L.tv(xp, eps) - L.tv(xm, eps)) / (2 * h)
        assert abs(fd - g[idx]) < 1e-7 * max(1.0, abs(fd))


def test_loss_grad_directional_fd():
    """C3's gradcheck geometry: 96^2 image, 3x3 patches, 9x9 PSFs."""
    sx, csf, sf = (96, 96), (3, 3), (9, 9)
    wa = O.Win2WinAux(sx, csf, 0.5)
    x, f, g = case_inputs(21, sx, csf, sf)
    y = O.forward(x, f, wa) + 0.1 * g
    lam, eps = 1e-3, 1e-2
    gx, gf = L.loss_grad(x, f, y, wa, lam, eps)
    gx0, _ = L.loss_grad(x, f, y, wa, 0.0, eps)
    rng = np.random.default_rng(22)
    for _ in range(3):
        vx = rng.standard_normal(x.shape)
        vf = rng.standard_normal(f.shape)
        h = 3e-6
        fd = (L.loss(x + h * vx, f + h * vf, y, wa, lam, eps) - L.loss(x - h * vx, f - h * vf, y, wa, lam, eps)) / (2 * h)
        an = np.sum(gx * vx) + np.sum(gf * vf)
        # the TV term carries ~1e-3 of the derivative: a tolerance of 1e-6 sees it
        assert abs(np.sum((gx - gx0) * vx)) > 1e-4 * abs(an)
        assert abs(fd - an) < 1e-6 * abs(an)
