"""FP64 NumPy restatement of the differentiable EFF/OLA loss
(test infrastructure only):

    L(x, f) = 1/2 |B(x, f) - y|^2 + lam * TV(x)

B is the forward operator of olaGPU.py:170-257 (ola_oracle.forward); the
gradients use its exact adjoints.  TV is isotropic and eps-smoothed,
TV(x) = sum_ij sqrt(dx_ij^2 + dy_ij^2 + eps^2) with forward differences that
are zero past the last column/row (the reference has no TV prior; this term
is defined here).
"""
import numpy as np

from . import ola_oracle as O


def tv(x, eps):
    x = np.asarray(x, dtype=np.float64)
    dx = np.zeros_like(x)
    dy = np.zeros_like(x)
    dx[..., :, :-1] = x[..., :, 1:] - x[..., :, :-1]
    dy[..., :-1, :] = x[..., 1:, :] - x[..., :-1, :]
    return float(np.sum(np.sqrt(dx * dx + dy * dy + eps * eps)))


def tv_grad(x, eps):
    x = np.asarray(x, dtype=np.float64)
    dx = np.zeros_like(x)
    dy = np.zeros_like(x)
    dx[..., :, :-1] = x[..., :, 1:] - x[..., :, :-1]
    dy[..., :-1, :] = x[..., 1:, :] - x[..., :-1, :]
    n = np.sqrt(dx * dx + dy * dy + eps * eps)
    a, b = dx / n, dy / n
    g = -(a + b)
    g[..., :, 1:] += a[..., :, :-1]
    g[..., 1:, :] += b[..., :-1, :]
    return g


def loss(x, f, y, winaux, lam, eps, mode='valid'):
    r = O.forward(x, f, winaux, mode) - y
    return 0.5 * float(np.sum(r * r)) + lam * tv(x, eps)


def loss_grad(x, f, y, winaux, lam, eps, mode='valid'):
    """(dL/dx, dL/df) with the exact adjoints."""
    x = np.asarray(x, dtype=np.float64)
    f = np.asarray(f, dtype=np.float64)
    r = O.forward(x, f, winaux, mode) - y
    gx = O.adjoint_x(r, f, winaux, x.shape, mode, exact=True) + lam * tv_grad(x, eps)
    gf = O.adjoint_f(r, x, winaux, f.shape[-2:], mode, exact=True)
    return gx, gf
