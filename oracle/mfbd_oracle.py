"""FP64 NumPy restatement of the online MFBD frame step (vmbd_chimney.py:110-147)
and of projected Barzilai-Borwein with a fixed iteration count
(optGPU.py:134-297).  Test infrastructure only.

The PBB start pair is given explicitly (the reference draws the second point
from curand, optGPU.py:258; the caller passes the same uniform draw u to the
device path and here), so the iterate sequence is deterministic.
"""
import numpy as np

from . import filters_oracle as FO
from . import ola_oracle as O


def pbb_fixed(grad, obj, z0, u, iters, unconstrained=False):
    """optGPU.py:256-297 start, then `iters` updates of optGPU.py:200-245.

    Start pair: z0 and y = u - (z0 + 0.5) (optGPU.py:258-259); the point with
    the lower objective is the iterate (ties go to y: 'if fx < fy').  Update
    `it` (1-based): dx = x - oldx, dg = g - oldg, both zeroed where x == 0 and
    g > 0 (clip2bound, optGPU.py:129-132); step = sum dx dg / (1e-5 + sum dg^2)
    on odd `it`, sum dx^2 / (1e-5 + sum dx dg) on even; x <- max(x - step g, 0).
    """
    z0 = np.asarray(z0, dtype=np.float64)
    y = np.asarray(u, dtype=np.float64) - (z0 + 0.5)
    if obj(z0) < obj(y):
        x, ox = z0, y
    else:
        x, ox = y, z0
    g, og = grad(x), grad(ox)
    for it in range(1, iters + 1):
        dx = (x - ox).ravel().copy()
        dg = (g - og).ravel().copy()
        if not unconstrained:
            m = (x.ravel() == 0) & (g.ravel() > 0)
            dx[m] = 0
            dg[m] = 0
        if it % 2 == 0:
            step = (dx @ dx) / (1e-5 + dx @ dg)
        else:
            step = (dx @ dg) / (1e-5 + dg @ dg)
        xn = x - step * g
        if not unconstrained:
            xn = np.maximum(xn, 0.0)
        ox, og = x, g
        x = xn
        g = grad(x)
    return x


def normalize(f):
    """gputools.py:1025-1044 (3-D): f_k / sum f_k where the sum exceeds 1e-6."""
    f = np.array(f, dtype=np.float64)
    for k in range(f.shape[0]):
        s = f[k].sum()
        if s > 1e-6:
            f[k] /= s
    return f


def rl_step(x, f, y, winaux, sf, tol=1e-10, exact=False):
    """vmbd_chimney.py:137-147: multiplicative x-step with the saturation mask."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    mask = (y < 1).astype(np.float64)
    F = O.OlaOracle(f, x.shape, 'valid', winaux)
    num = F.cnvtp(mask * y, exact=exact)
    den = F.cnvtp(mask * F.cnv(x), exact=exact)
    fac = np.maximum(num / (den + tol), tol)
    x = x * fac
    xmax = x[sf[0]:-sf[0], sf[1]:-sf[1]].max()
    return np.minimum(x, xmax)


def f_step_pbb(x, y, winaux, sf, u, iters, beta, exact=False, f0=None):
    """PSF estimate: PBB on 1/2 |X f - y|^2 + beta/2 |f|^2 over f >= 0 from
    f0 (X'y when None, olaGPU.py:388-392), then normalize."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    X = O.OlaOracle(x, sf, 'valid', winaux)

    def obj(z):
        r = X.cnv(z) - y
        return 0.5 * np.sum(r * r) + 0.5 * beta * np.sum(z * z)

    def grad(z):
        return X.cnvtp(X.cnv(z) - y, exact=exact) + beta * z

    z0 = X.cnvtp(y, exact=exact) if f0 is None else np.asarray(f0, dtype=np.float64)
    return normalize(pbb_fixed(grad, obj, z0, u, iters))


def x_step_pbb(x, f, y, winaux, u, iters, exact=False):
    """Latent image: PBB on 1/2 |mask (F x - y)|^2 over x >= 0 from x."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    mask = (y < 1).astype(np.float64)
    F = O.OlaOracle(f, x.shape, 'valid', winaux)

    def obj(z):
        r = mask * (F.cnv(z) - y)
        return 0.5 * np.sum(r * r)

    def grad(z):
        return F.cnvtp(mask * (F.cnv(z) - y), exact=exact)

    return pbb_fixed(grad, obj, x, u, iters)


def frame_step(x, y, winaux, sf, fu, f_iters, beta, tol=1e-10, exact=False, x_solver='rl', xu=None, x_iters=0,
               f0=None, prior=False):
    """One frame of the online loop with the device f-solver (PBB) started
    from f0 (X'y when None); `prior`: bilateral (5, 3, 0.1) then three shock
    steps (dt 0.1) on the latent image."""
    f = f_step_pbb(x, y, winaux, sf, fu, f_iters, beta, exact, f0)
    if x_solver == 'rl':
        x = rl_step(x, f, y, winaux, sf, tol, exact)
    else:
        x = x_step_pbb(x, f, y, winaux, xu, x_iters, exact)
    if prior:
        x = FO.shock_filter(FO.bilateral(x, 5, 3.0, 0.1), 3, 0.1, 1.0)
    return x, f
