"""Drop-in of the reference's optGPU.py: projected Barzilai-Borwein (PBB)
for min f(x) s.t. x >= 0 (optGPU.py:134-358), on the device.

The BB statistics (sum dx^2, sum dx dg, sum dg^2 under the clip2bound mask)
and the projected step x <- max(x - step g, 0) are two fused kernels; the
step length never leaves the device.  ``PBB`` keeps the reference's control
flow (termination on the objective change needs the objective on the host,
optGPU.py:326-330); ``pbb_fixed`` runs a fixed number of iterations with no
host synchronisation, so a caller can capture it in a HIP graph.

Deterministic start: the reference draws the second starting point from an
unseeded generator (optGPU.py:258); here it comes from a seeded one
(``Solopt.seed``), with the same formula rand - (x0 + 0.5).
"""
import time

import numpy as np
import torch

from . import _lib
from . import gputools


class Solopt:
    """optGPU.py:56-127 defaults."""

    def __init__(self):
        self.maxiter = 200
        self.maxtime = 100
        self.maxmem = 10
        self.time_limit = 0
        self.use_tolx = 0
        self.use_tolo = 1
        self.use_tolg = 0
        self.use_kkt = 0
        self.tolx = 1e-8
        self.tolo = 1e-6
        self.tolk = 1e-8
        self.tolg = 1e-8
        self.verbose = 0
        self.tau = 1e-5
        self.compute_obj = 1
        self.compute_both = 0
        self.asgui = 0
        self.max_func_evals = 10
        self.pbb_gradient_norm = 1e-9
        self.beta = 0.0498
        self.sigma = 0.298
        self.unconstrained = False
        self.seed = 120


class MVM_Objective:
    """optGPU.py:19-31: f(x) = 1/2 |A x - b|^2 (dense test objective)."""

    def __init__(self, A, b):
        self.A = A
        self.b = b

    def compute_obj(self, x):
        r = self.A @ x - self.b
        return 0.5 * gputools.dot(r, r)

    def compute_grad(self, x):
        return (self.A.T @ (self.A @ x - self.b)).contiguous()


class CNV_Objective:
    """optGPU.py:33-46: f(z) = 1/2 |X z - y|^2 with an OlaGPU instance X.
    exact=True uses the true transpose for the gradient."""

    def __init__(self, X, y, exact=True, beta=0.0):
        self.X = X
        self.y = gputools._dev(y)
        self.exact = exact
        self.beta = float(beta)

    def _res(self, z):
        return gputools.ew(gputools.EW_SUB, self.X.cnv(z), self.y)

    def compute_obj(self, z):
        r = self._res(z)
        o = 0.5 * gputools.dot(r, r)
        if self.beta > 0:
            o += 0.5 * self.beta * gputools.dot(z, z)
        return o

    def compute_grad(self, z):
        g = self.X.cnvtp(self._res(z), exact=self.exact)
        if self.beta > 0:
            g = gputools.ew(gputools.EW_AXPY, g, z, self.beta)
        return g

    def compute_both(self, z):
        r = self._res(z)
        o = 0.5 * gputools.dot(r, r)
        g = self.X.cnvtp(r, exact=self.exact)
        if self.beta > 0:
            o += 0.5 * self.beta * gputools.dot(z, z)
            g = gputools.ew(gputools.EW_AXPY, g, z, self.beta)
        return o, g


class Objective:
    """optGPU.py:48-53."""

    def __init__(self, obj=None, grad=None, both=None):
        self.compute_obj = obj
        self.compute_grad = grad
        self.compute_both = both


class _Scratch:
    def __init__(self, n, device):
        nb = _lib.lib().vmbd_reduce_blocks(n)
        self.partials = torch.empty(3 * nb, dtype=torch.float64, device=device)
        self.step = torch.empty(1, dtype=torch.float32, device=device)


def pbb_step(x, oldx, g, oldg, it, out, scratch, unconstrained=False):
    """One projected BB update: out = P(x - step g) with the BB1/BB2 step of
    iteration `it` (1-based) from dx = x - oldx, dg = g - oldg."""
    n = x.numel()
    _lib.call('vmbd_pbb_stats', _lib.ptr(x), _lib.ptr(oldx), _lib.ptr(g), _lib.ptr(oldg), n, int(unconstrained),
              _lib.ptr(scratch.partials), None, _lib.stream_ptr())
    _lib.call('vmbd_pbb_update', _lib.ptr(out), _lib.ptr(x), _lib.ptr(g), n, _lib.ptr(scratch.partials), int(it),
              int(unconstrained), _lib.ptr(scratch.step), _lib.stream_ptr())
    _lib.wrote(out)
    return out


class PBB:
    """optGPU.py:134-358.  Result in ``.result``; ``.iter``, ``.status``,
    ``.term_reason``, ``.obj``."""

    def __init__(self, objective, x_init, options):
        self.objective = objective
        self.options = options
        self.time_start = time.time()
        self.iter = 0
        self.status = 'Failure'
        self.term_reason = 0
        self.initialisation(gputools._dev(x_init))
        if np.sqrt(gputools.dot(self.x, self.x)) < 1e-12:                 # optGPU.py:185-186
            raise IOError('Initial vector close to zero. Cannot proceed')
        sc = _Scratch(self.x.numel(), self.x.device)
        for _ in range(options.maxiter):
            self.iter += 1
            self.check_termination()
            if self.term_reason:
                break
            xnew = torch.empty_like(self.x)
            pbb_step(self.x, self.oldx, self.g, self.oldg, self.iter, xnew, sc, options.unconstrained)
            self.oldx, self.oldg = self.x, self.g
            self.x = xnew
            if options.compute_both and objective.compute_both is not None:
                self.oldobj = self.obj
                self.obj, self.g = objective.compute_both(self.x)
            elif options.compute_obj:
                self.g = objective.compute_grad(self.x)
                self.oldobj = self.obj
                self.obj = objective.compute_obj(self.x)
            else:
                self.g = objective.compute_grad(self.x)
        self.time = time.time() - self.time_start
        self.status = 'Success'
        self.result = self.x

    def initialisation(self, x_init):
        """optGPU.py:256-297 (seeded second point)."""
        gen = torch.Generator(device=x_init.device)
        gen.manual_seed(int(getattr(self.options, 'seed', 120)))
        y = torch.rand(x_init.shape, generator=gen, device=x_init.device, dtype=torch.float32)
        y = gputools.ew(gputools.EW_SUB, y, x_init)
        y = gputools.ew(gputools.EW_AXPY, y, torch.ones_like(y), -0.5)
        ob = self.objective
        if self.options.compute_both and ob.compute_both is not None:
            fx, gx = ob.compute_both(x_init)
            fy, gy = ob.compute_both(y)
        else:
            fx, fy = ob.compute_obj(x_init), ob.compute_obj(y)
            gx = gy = None
        if fx < fy:
            self.x, self.oldx, self.obj, self.oldobj = x_init, y, fx, fy
            self.g = gx if gx is not None else ob.compute_grad(x_init)
            self.oldg = gy if gy is not None else ob.compute_grad(y)
        else:
            self.x, self.oldx, self.obj, self.oldobj = y, x_init, fy, fx
            self.g = gy if gy is not None else ob.compute_grad(y)
            self.oldg = gx if gx is not None else ob.compute_grad(x_init)

    def check_termination(self):
        """optGPU.py:306-358 (time limit, tolo after iteration 2, maxiter)."""
        o = self.options
        if o.time_limit and time.time() - self.time_start >= o.maxtime:
            self.term_reason = 'Exceeded time limit'
            return
        if o.use_tolo and self.iter > 2:
            if abs(self.obj - self.oldobj) < o.tolo:
                self.term_reason = 'Relative change in objvalue small enough'
                return
        if self.iter >= o.maxiter:
            self.term_reason = 'Maximum number of iterations reached'
            return
        self.term_reason = 0


def pbb_fixed(grad_fn, x, oldx, g, oldg, iters, unconstrained=False, scratch=None, bufs=None):
    """`iters` PBB iterations with no host synchronisation.

    grad_fn(x, out) writes the gradient at x into `out`.  x, oldx, g, oldg
    are device buffers; the iterate moves between two fixed buffer pairs, so
    the sequence of launches is static and can be captured in a HIP graph.
    Returns (x, oldx, g, oldg) after the last iteration.
    """
    sc = scratch or _Scratch(x.numel(), x.device)
    for it in range(1, iters + 1):
        pbb_step(x, oldx, g, oldg, it, oldx, sc, unconstrained)   # new iterate overwrites oldx
        x, oldx = oldx, x
        grad_fn(x, oldg)                                          # new gradient overwrites oldg
        g, oldg = oldg, g
    return x, oldx, g, oldg
