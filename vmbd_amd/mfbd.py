"""Online multi-frame blind deconvolution (vmbd_chimney.py:89-147) on the
EFF/OLA plan.

Per observed frame y (vmbd_chimney.py:110-147):

    mask = (y < 1)                                          saturation mask
    f    = argmin_{f >= 0} 1/2 |y - X f|^2 + beta/2 |f|^2   X: current latent x
    f_k  = f_k / sum(f_k)   when sum(f_k) > 1e-6            gputools.normalize
    x   <- x * max(F'(mask y) / (F'(mask F x) + tol), tol)
    x   <- min(x, max of x away from an sf-wide border)

Solvers:

* ``f_solver='lbfgsb'``: host L-BFGS-B through ``OlaGPU.deconv`` (the
  reference's, started from X'y, up to ``optiter`` evaluations);
* ``f_solver='pbb'``: projected Barzilai-Borwein on the device for a fixed
  ``f_iters`` iterations, started like optGPU.PBB (optGPU.py:256-297) from a
  first point and a seeded random second point.  The first point is the
  previous frame's PSFs (``f_init='warm'``, flat 1/(sf0 sf1) PSFs on the
  first frame, the reference's commented-out alternative at olaGPU.py:393)
  or X'y (``f_init='xty'``, the L-BFGS-B start of olaGPU.py:388-392, whose
  scale (~1e6 per PSF on 8-bit frames) a few PBB iterations do not undo);
* ``x_solver='rl'``: the multiplicative step above (the reference's);
* ``x_solver='pbb'``: ``x_iters`` PBB iterations on 1/2 |mask (F x - y)|^2
  over x >= 0;
* ``prior='bilateral_shock'``: after the x-step, a bilateral filter
  (half width 5, sigma_d 3, sigma_r 0.1, gputools.py:251-285) and three
  shock-filter steps (dt 0.1, gputools.py:1282-1314) on the latent image.

With ``f_solver='pbb'`` a frame is a fixed sequence of launches with no host
synchronisation; ``capture=True`` records it once in a HIP graph and replays
it for every frame (the random second start points are drawn before each
replay).  The latent image and the PSFs stay on the device between frames.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import gputools
from .imagetools import win2winaux
from .ola import OlaPlan
from .olaGPU import OlaGPU
from .optGPU import _Scratch, pbb_step


def _pair(v):
    a = np.asarray(v, dtype=np.int64).ravel()
    return (int(a[0]), int(a[0])) if a.size == 1 else (int(a[0]), int(a[1]))


def initial_latent(frames, N0=20):
    """vmbd_chimney.py:89-97: the first N0-1 frames (y_0001 .. y_0019 for
    N0 = 20; frames[0] is y_0001) summed and divided by N0."""
    y0 = 0.
    for i in range(N0 - 1):
        y0 = y0 + np.asarray(frames[i], dtype=np.float32)
    return (y0 / N0).astype(np.float32)


class OnlineMFBD:
    """State of the online loop: latent image ``x`` (sy + sf - 1), PSFs ``f``
    (P, sf0, sf1), and the plan shared by every frame.

        m = OnlineMFBD(y0, sf=(20, 20), csf=(3, 3))
        for y in frames:
            x, f = m.step(y)
    """

    def __init__(self, y0, sf, csf=(3, 3), overlap=0.5, f_solver='lbfgsb', x_solver='rl', f_alpha=0.,
                 f_beta=0.1, optiter=500, tol=1e-10, f_iters=20, x_iters=20, exact=False, seed=120, capture=False,
                 f_init='warm', prior=None, device=None, tune=False):
        if f_solver not in ('lbfgsb', 'pbb'):
            raise ValueError("f_solver must be 'lbfgsb' or 'pbb'")
        if x_solver not in ('rl', 'pbb'):
            raise ValueError("x_solver must be 'rl' or 'pbb'")
        if capture and f_solver != 'pbb':
            raise ValueError('capture needs the device f-step (f_solver=pbb)')
        if f_init not in ('warm', 'xty'):
            raise ValueError("f_init must be 'warm' or 'xty'")
        self.f_init = f_init
        if prior not in (None, 'bilateral_shock'):
            raise ValueError("prior must be None or 'bilateral_shock'")
        self.prior = prior
        self.device = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
        y0 = gputools._dev(y0)
        self.sy = tuple(int(v) for v in y0.shape)
        self.sf = _pair(sf)
        self.sx = (self.sy[0] + self.sf[0] - 1, self.sy[1] + self.sf[1] - 1)
        self.csf = _pair(csf)
        self.winaux = win2winaux(self.sx, self.csf, overlap, device=self.device)
        self.P = self.csf[0] * self.csf[1]
        self.f_solver, self.x_solver = f_solver, x_solver
        self.f_alpha, self.f_beta = float(f_alpha), float(f_beta)
        if f_solver == 'pbb' and self.f_alpha != 0.:
            raise NotImplementedError('the device f-step has no gradient (alpha) term')
        self.optiter, self.tol = int(optiter), float(tol)
        self.f_iters, self.x_iters = int(f_iters), int(x_iters)
        self.exact = bool(exact)
        self.capture = bool(capture)
        self.gen = torch.Generator(device=self.device)
        self.gen.manual_seed(int(seed))
        # the FFT tile and launch shapes of the plan: the smallest tile and the
        # library's own shapes, or (tune=True) the autotuner's fastest
        self.tuning = dict(tile=None, ola_threads=0, col_threads=0)
        if tune:
            from .autotune import best_config
            self.tuning, _ = best_config(self.sx, self.sf, self.winaux, frames=1, overlap=overlap,
                                         device=self.device, workload='cached')
        self.plan = OlaPlan(self.sx, self.sf, self.winaux, 'valid', tile=self.tuning['tile'], device=self.device)
        self.plan.set_tuning(self.tuning['ola_threads'], self.tuning['col_threads'])
        # vmbd_chimney.py:97: x = impad(y0, sf - 1)
        self.x = gputools.impad_gpu(y0, (self.sf[0] - 1, self.sf[1] - 1))
        self.f = None
        self.frame = 0
        self._alloc()
        self._graph = None

    # ------------------------------------------------------------ buffers
    def _alloc(self):
        d, f32 = self.device, torch.float32
        fshape = (self.P,) + self.sf
        spec_shape = (self.P, self.plan.tile[1] // 2 + 1, self.plan.tile[0])
        self.y = torch.zeros(self.sy, dtype=f32, device=d)
        self.mask = torch.zeros(self.sy, dtype=f32, device=d)
        self.ym = torch.zeros(self.sy, dtype=f32, device=d)
        self.r = torch.zeros(self.sy, dtype=f32, device=d)
        self.specX = torch.zeros(spec_shape, dtype=torch.complex64, device=d)
        self.specF = torch.zeros(spec_shape, dtype=torch.complex64, device=d)
        self.dots = torch.zeros(4, dtype=torch.float64, device=d)
        nf = int(np.prod(fshape))
        nx = self.sx[0] * self.sx[1]
        self.sc_f = _Scratch(nf, d)
        self.sc_x = _Scratch(nx, d)
        self.dot_part = torch.zeros(3 * _lib.lib().vmbd_reduce_blocks(max(nf, nx, self.sy[0] * self.sy[1])),
                                    dtype=torch.float64, device=d)
        self.fz = [torch.zeros(fshape, dtype=f32, device=d) for _ in range(4)]     # z, oldz, g, oldg
        self.fu = torch.zeros(fshape, dtype=f32, device=d)
        self.xz = [torch.zeros(self.sx, dtype=f32, device=d) for _ in range(3)]    # oldx, g, oldg
        self.xu = torch.zeros(self.sx, dtype=f32, device=d)
        self.tmp = [torch.zeros(self.sx, dtype=f32, device=d) for _ in range(2)]
        self.num = torch.zeros(self.sx, dtype=f32, device=d)
        self.den = torch.zeros(self.sx, dtype=f32, device=d)
        self.part_max = torch.zeros(_lib.lib().vmbd_reduce_blocks(nx), dtype=f32, device=d)
        self.xmax = torch.zeros(1, dtype=f32, device=d)
        # PSF state: flat PSFs until the first frame has been processed
        self.f_dev = torch.full(fshape, 1.0 / (self.sf[0] * self.sf[1]), dtype=f32, device=d)

    # ---------------------------------------------------------- primitives
    def _ew(self, op, out, a, b=None, s=0.0):
        _lib.call('vmbd_ew', _lib.ptr(out), _lib.ptr(a), _lib.ptr(b), float(s), a.numel(), int(op),
                  _lib.stream_ptr())

    def _dot(self, a, b, slot):
        out = ctypes.c_void_p(self.dots.data_ptr() + slot * self.dots.element_size())
        _lib.call('vmbd_dot', _lib.ptr(a), _lib.ptr(b), a.numel(), _lib.ptr(self.dot_part), out, _lib.stream_ptr())

    def _grad_f(self, z, out, dots_at=None):
        """out = X'(X z - y) + beta z (X: cached image spectra); the residual
        and the Tikhonov term are formed in the operators' final stores."""
        p = self.plan
        p.forward_cached('X', self.specX, z, out=self.r, sub=self.y)
        p.adjoint_cached('X', self.specX, self.r, exact=self.exact, out=out,
                         z=z if self.f_beta != 0. else None, beta=self.f_beta)
        if dots_at is not None:
            self._dot(self.r, self.r, dots_at)
            self._dot(z, z, dots_at + 1)

    def _grad_x(self, z, out, dots_at=None):
        """out = F'(mask (F z - y)) (F: cached PSF spectra)."""
        p = self.plan
        p.forward_cached('F', self.specF, z, out=self.r, sub=self.y, mul=self.mask)
        p.adjoint_cached('F', self.specF, self.r, exact=self.exact, out=out)
        if dots_at is not None:
            self._dot(self.r, self.r, dots_at)
            self._dot(z, z, dots_at + 1)

    def _pbb(self, grad, z, oldz, g, oldg, u, iters, beta, sc):
        """optGPU.PBB with a fixed iteration count: start pair (z, u - (z + 0.5)),
        the better one becomes the iterate (on the device), then `iters`
        projected BB steps.  Returns the buffer holding the final iterate."""
        self._ew(gputools.EW_SUB, oldz, u, z)
        self._ew(gputools.EW_ADDS, oldz, oldz, None, -0.5)
        grad(z, g, dots_at=0)
        grad(oldz, oldg, dots_at=2)
        _lib.call('vmbd_pbb_select', _lib.ptr(z), _lib.ptr(oldz), _lib.ptr(g), _lib.ptr(oldg), z.numel(),
                  _lib.ptr(self.dots), float(beta), _lib.stream_ptr())
        for it in range(1, iters + 1):
            pbb_step(z, oldz, g, oldg, it, oldz, sc)      # new iterate overwrites oldz
            z, oldz = oldz, z
            grad(z, oldg)                                # new gradient overwrites oldg
            g, oldg = oldg, g
        return z

    # ------------------------------------------------------------- f-step
    def _f_step_device(self):
        p = self.plan
        p.spectra('X', self.x, out=self.specX)
        z, oldz, g, oldg = self.fz
        if self.f_init == 'warm':
            z.copy_(self.f_dev)
        else:   # X'y (olaGPU.py:388-392)
            p.adjoint_cached('X', self.specX, self.y, exact=self.exact, out=z)
        res = self._pbb(self._grad_f, z, oldz, g, oldg, self.fu, self.f_iters, self.f_beta, self.sc_f)
        self.f_dev.copy_(res)
        _lib.call('vmbd_psf_normalize', _lib.ptr(self.f_dev), self.P, self.sf[0] * self.sf[1], 1e-6,
                  _lib.stream_ptr())

    def _f_step_lbfgsb(self):
        X = OlaGPU(self.x, self.sf, 'valid', winaux=self.winaux, exact=self.exact)
        f = X.deconv(self.y, mode='lbfgsb', alpha=self.f_alpha, beta=self.f_beta, maxfun=self.optiter,
                     verbose=-1)[0]
        f = gputools.normalize(np.asarray(f, dtype=np.float64))           # host, as vmbd_chimney.py:133
        self.f_dev.copy_(torch.as_tensor(f.astype(np.float32)))

    # ------------------------------------------------------------- x-step
    def _x_step(self):
        p = self.plan
        p.spectra('F', self.f_dev, out=self.specF)
        if self.x_solver == 'rl':
            p.forward_cached('F', self.specF, self.x, out=self.r, mul=self.mask)
            p.adjoint_cached('F', self.specF, self.ym, exact=self.exact, out=self.num)
            p.adjoint_cached('F', self.specF, self.r, exact=self.exact, out=self.den)
            _lib.call('vmbd_rl_update', _lib.ptr(self.x), _lib.ptr(self.num), _lib.ptr(self.den), self.sx[0],
                      self.sx[1], self.tol, self.sf[0], self.sf[1], _lib.ptr(self.part_max), _lib.ptr(self.xmax),
                      _lib.stream_ptr())
        else:
            oldz, g, oldg = self.xz
            res = self._pbb(self._grad_x, self.x, oldz, g, oldg, self.xu, self.x_iters, 0.0, self.sc_x)
            if res is not self.x:
                self.x.copy_(res)

    def _prior(self):
        """Bilateral filter, then three shock steps, ping-ponging between two
        buffers so that the result lands in x (no allocation)."""
        H, W = self.sx
        st = _lib.stream_ptr()
        t1, t2 = self.tmp
        _lib.call('vmbd_bilateral', _lib.ptr(t1), _lib.ptr(self.x), H, W, 5, 3.0, 0.1, st)
        for src, dst in ((t1, t2), (t2, t1), (t1, self.x)):
            _lib.call('vmbd_shock_step', _lib.ptr(dst), _lib.ptr(src), H, W, 0.1, 1.0, st)

    def _frame_body(self):
        # vmbd_chimney.py:117-118: mask = 1. * (y < 1)
        self._ew(gputools.EW_LT, self.mask, self.y, None, 1.0)
        self._ew(gputools.EW_MUL, self.ym, self.mask, self.y)
        if self.f_solver == 'pbb':
            self._f_step_device()
        else:
            self._f_step_lbfgsb()
        self._x_step()
        if self.prior:
            self._prior()

    def _draw(self):
        torch.rand(self.fu.shape, generator=self.gen, device=self.device, out=self.fu)
        if self.x_solver == 'pbb':
            torch.rand(self.xu.shape, generator=self.gen, device=self.device, out=self.xu)

    # --------------------------------------------------------------- step
    def step(self, y):
        """Process one observed frame; returns (x, f) device tensors (views of
        the loop's state: copy them to keep a frame's result)."""
        y = gputools._dev(y)
        if tuple(y.shape) != self.sy:
            raise IOError('Sizes incompatible. Expected blurred image!')
        self.y.copy_(y)
        self._draw()
        if not self.capture:
            self._frame_body()
        else:
            if self._graph is None:
                self._record()
            self._graph.replay()
        # the latent image and the PSFs were written through raw pointers
        # (x-step, prior, normalisation): mark them modified
        _lib.wrote(self.x, self.f_dev)
        self.frame += 1
        self.f = self.f_dev
        return self.x, self.f_dev

    def _record(self):
        # one eager frame on a side stream initialises every lazy buffer, then
        # the same launches are recorded; the latent image and the PSF state
        # are restored so the recorded frame is the first one processed
        x0 = self.x.clone()
        f0 = self.f_dev.clone()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._frame_body()
        torch.cuda.current_stream().wait_stream(s)
        self.x.copy_(x0)
        self.f_dev.copy_(f0)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._frame_body()
        self._graph = g
