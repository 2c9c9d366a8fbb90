"""Operator plans over torch tensors: the EFF/OLA forward and its transposes
on the HIP kernels of libvmbd_hip.so.

    plan = OlaPlan(sx, sf, winaux, mode='valid', max_frames=1)
    y    = plan.forward(x, f)             # B(x,f) = F x = X f
    xbar = plan.adjoint_x(g, f)           # F' g
    fbar = plan.adjoint_f(g, x)           # X' g
    xbar, fbar = plan.adjoint_xf(g, x, f)

Tensors are float32 on the plan's device.  A single frame is x (H, W),
f (P, sf0, sf1), g/y (sy0, sy1); a batch of B <= max_frames frames adds a
leading dimension to each and runs all patches of all frames in the same
launches.  Work is enqueued on the current torch stream.  ``exact=True`` is
the true adjoint, ``exact=False`` the reference's patch-local transpose
(olaGPU.py:276-311).

``adjoint_xf`` reuses the row transforms of x and f computed by the
preceding ``forward`` when it is given the very same tensors (same storage,
unmodified since, same stream, nothing else run on the plan in between).
"Unmodified" is torch's version counter: torch in-place ops move it, and
so does every library wrapper that writes a tensor through a raw pointer
(``_lib.wrote``).  A write that bypasses both (a raw pointer of the
caller's own) must pass ``reuse=False``.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .imagetools import fft_tile


def _as_dev(t, device, name):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f'{name} must be a torch.Tensor')
    if t.device != device:
        raise ValueError(f'{name} is on {t.device}, plan is on {device}')
    if t.dtype != torch.float32:
        raise TypeError(f'{name} must be float32, got {t.dtype}')
    return t.contiguous()


class OlaPlan:
    def __init__(self, sx, sf, winaux, mode='valid', tile=None, device=None, max_frames=1, separable=True):
        if mode not in _lib.MODES:
            raise NotImplementedError('Not a valid mode!')
        sx = tuple(int(v) for v in sx)
        sf = tuple(int(v) for v in sf)
        if tuple(int(v) for v in winaux.sx) != sx:
            raise IOError('Size missmatch between winaux and image size!')
        self.device = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
        self.sx, self.sf, self.mode = sx, sf, mode
        self.winaux = winaux
        self.csf = winaux.csf
        self.P = winaux.csf[0] * winaux.csf[1]
        self.sw = tuple(int(v) for v in winaux.sw)
        self.nhop = tuple(int(v) for v in winaux.nhop)
        self.max_frames = int(max_frames)
        self.tile = tuple(int(v) for v in (tile if tile is not None else fft_tile(self.sw, sf)))
        g = _lib.VmbdGeom(sx[0], sx[1], sf[0], sf[1], self.csf[0], self.csf[1], self.sw[0], self.sw[1],
                          self.nhop[0], self.nhop[1], self.tile[0], self.tile[1], _lib.MODES[mode], self.max_frames)
        table = np.ascontiguousarray(winaux.win_table, dtype=np.float32)
        idx = np.ascontiguousarray(winaux.win_idx, dtype=np.int32)
        L = _lib.lib()
        with torch.cuda.device(self.device):
            h = L.vmbd_plan_create(ctypes.byref(g), table.ctypes.data_as(ctypes.c_void_p), int(table.shape[0]),
                                   idx.ctypes.data_as(ctypes.c_void_p))
        if not h:
            raise _lib.VmbdError(f'vmbd_plan_create: {_lib.last_error()}', L.vmbd_last_status())
        self._h = ctypes.c_void_p(h)
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(L.vmbd_plan_out_shape(self._h, ctypes.byref(a), ctypes.byref(b)), 'out_shape')
        self.sy = (a.value, b.value)
        self.spec_elems = int(L.vmbd_plan_spectrum_elems(self._h))
        self.workspace_bytes = int(L.vmbd_plan_workspace_bytes(self._h))
        self._rows_key = None
        # separable windows: the x-adjoint's overlap-add in summed form (the
        # library checks the factors against the tables and refuses otherwise)
        self.separable = False
        if separable and hasattr(winaux, 'window_factors'):
            fa, fb, floor = winaux.window_factors()
            with torch.cuda.device(self.device):
                rc = L.vmbd_plan_set_window_factors(self._h, fa.ctypes.data_as(ctypes.c_void_p),
                                                    fb.ctypes.data_as(ctypes.c_void_p), float(floor))
            if rc == 0:
                self.separable = True
                self.workspace_bytes = int(L.vmbd_plan_workspace_bytes(self._h))
            elif rc != _lib.VMBD_E_UNSUP:
                _lib.check(rc, 'vmbd_plan_set_window_factors')

    def image_rows_shared(self, nframes=1):
        """True when the forward and adjoint_xf of `nframes` frames transform
        each image row once per patch column (separable windows, and by
        default launches large enough to fill the GPU; a single 1024^2 frame
        is not), and adjoint_f follows the same choice.  The calls that transform the
        image rows alone (the image spectra, forward_cached with PSF spectra)
        share them whenever the windows are separable and image_rows is not
        2."""
        rc = _lib.lib().vmbd_plan_image_rows_shared(self._h, int(nframes))
        if rc < 0:
            _lib.check(rc, 'vmbd_plan_image_rows_shared')
        return rc == 1

    # -- launch shapes --------------------------------------------------------
    def set_tuning(self, ola_threads=0, col_threads=0, image_rows=0):
        """Threads per workgroup of the overlap-add passes (128 / 192 / 256)
        and of the column passes (128 / 256); 0 = automatic.  Results are
        the same bits for every setting of these two.  image_rows (separable
        windows only): 1 transforms each image row once per patch column,
        2 once per patch, 0 chooses by launch size; the two forms round
        differently, each within the operators' accuracy."""
        t = _lib.VmbdTuning(int(ola_threads), int(col_threads), int(image_rows))
        _lib.check(_lib.lib().vmbd_plan_set_tuning(self._h, ctypes.byref(t)), 'set_tuning')
        self._rows_key = None  # the library dropped the forward's row transforms

    @property
    def tuning(self):
        t = _lib.VmbdTuning()
        _lib.check(_lib.lib().vmbd_plan_get_tuning(self._h, ctypes.byref(t)), 'get_tuning')
        return dict(ola_threads=t.ola_threads, col_threads=t.col_threads, image_rows=t.image_rows)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h:
            try:
                torch.cuda.synchronize(self.device)
                _lib.lib().vmbd_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    # -- shape checks -------------------------------------------------------
    def _frames(self, t, base_ndim, name):
        """Number of frames of t (1 for an unbatched tensor)."""
        if t.dim() == base_ndim:
            return 1, False
        if t.dim() == base_ndim + 1:
            nf = int(t.shape[0])
            if not 1 <= nf <= self.max_frames:
                raise ValueError(f'{name}: {nf} frames, plan allows 1..{self.max_frames}')
            return nf, True
        raise IOError('Size missmatch')

    def _x(self, x):
        x = _as_dev(x, self.device, 'x')
        nf, bat = self._frames(x, 2, 'x')
        if tuple(x.shape[-2:]) != self.sx:
            raise IOError('Size missmatch between winaux and image size!')
        return x, nf, bat

    def _f(self, f):
        f = _as_dev(f, self.device, 'f')
        nf, bat = self._frames(f, 3, 'f')
        if f.shape[-3] != self.P:
            raise IOError('Size missmatch between winaux and PSF size!')
        if tuple(f.shape[-2:]) != self.sf:
            raise IOError('Size missmatch')
        return f, nf, bat

    def _g(self, g):
        g = _as_dev(g, self.device, 'g')
        nf, bat = self._frames(g, 2, 'g')
        if tuple(g.shape[-2:]) != self.sy:
            raise IOError('Sizes incompatible. Expected blurred image!')
        return g, nf, bat

    @staticmethod
    def _same(*nfs):
        if len(set(n for n, _ in nfs)) != 1 or len(set(b for _, b in nfs)) != 1:
            raise ValueError('operands disagree on the number of frames')
        return nfs[0]

    def _empty(self, shape, out):
        if out is None:
            return torch.empty(shape, dtype=torch.float32, device=self.device)
        if (tuple(out.shape) != tuple(shape) or out.dtype != torch.float32 or not out.is_contiguous()
                or out.device != self.device):
            raise ValueError('out has the wrong shape/dtype/layout/device')
        _lib.wrote(out)  # the kernels write it through a raw pointer
        return out

    @staticmethod
    def _lead(nf, bat):
        return (nf,) if bat else ()

    def _key(self, x, f, nf):
        # the tensors themselves are held (not their addresses), so a freed
        # and reallocated buffer can never match a stale key; inference
        # tensors carry no version counter and are never reused
        if x.is_inference() or f.is_inference():
            return None
        return (x, x._version, f, f._version, nf, torch.cuda.current_stream().cuda_stream)

    def _key_matches(self, x, f, nf):
        k = self._rows_key
        if k is None or x.is_inference() or f.is_inference():
            return False
        kx, vx, kf, vf, knf, kst = k
        return (kx is x and kf is f and vx == x._version and vf == f._version and knf == nf
                and kst == torch.cuda.current_stream().cuda_stream)

    # -- operators ----------------------------------------------------------
    def forward(self, x, f, out=None):
        (x, nx, bx), (f, nfr, bf) = self._x(x), self._f(f)
        nf, bat = self._same((nx, bx), (nfr, bf))
        y = self._empty(self._lead(nf, bat) + self.sy, out)
        self._rows_key = None
        _lib.check(_lib.lib().vmbd_ola_forward(self._h, nf, _lib.ptr(x), _lib.ptr(f), _lib.ptr(y),
                                               _lib.stream_ptr()), 'forward')
        self._rows_key = self._key(x, f, nf)
        return y

    def adjoint_x(self, g, f, exact=True, out=None):
        (g, ng, bg), (f, nfr, bf) = self._g(g), self._f(f)
        nf, bat = self._same((ng, bg), (nfr, bf))
        xb = self._empty(self._lead(nf, bat) + self.sx, out)
        self._rows_key = None
        _lib.check(_lib.lib().vmbd_ola_adjoint_x(self._h, nf, _lib.ptr(g), _lib.ptr(f), _lib.ptr(xb),
                                                 int(bool(exact)), _lib.stream_ptr()), 'adjoint_x')
        return xb

    def adjoint_f(self, g, x, exact=True, out=None):
        (g, ng, bg), (x, nx, bx) = self._g(g), self._x(x)
        nf, bat = self._same((ng, bg), (nx, bx))
        fb = self._empty(self._lead(nf, bat) + (self.P,) + self.sf, out)
        self._rows_key = None
        _lib.check(_lib.lib().vmbd_ola_adjoint_f(self._h, nf, _lib.ptr(g), _lib.ptr(x), _lib.ptr(fb),
                                                 int(bool(exact)), _lib.stream_ptr()), 'adjoint_f')
        return fb

    def adjoint_xf(self, g, x, f, exact=True, out_x=None, out_f=None, reuse='auto'):
        (g, ng, bg), (x, nx, bx), (f, nfr, bf) = self._g(g), self._x(x), self._f(f)
        nf, bat = self._same((ng, bg), (nx, bx), (nfr, bf))
        xb = self._empty(self._lead(nf, bat) + self.sx, out_x)
        fb = self._empty(self._lead(nf, bat) + (self.P,) + self.sf, out_f)
        flags = 0
        if reuse == 'auto':
            reuse = self._key_matches(x, f, nf)
        if reuse:
            flags |= _lib.REUSE_ROWS
        if not reuse:
            self._rows_key = None
        _lib.check(_lib.lib().vmbd_ola_adjoint_xf(self._h, nf, _lib.ptr(g), _lib.ptr(x), _lib.ptr(f), _lib.ptr(xb),
                                                  _lib.ptr(fb), int(bool(exact)), flags, _lib.stream_ptr()),
                   'adjoint_xf')
        # the workspaces now hold the row transforms of this x and f
        self._rows_key = self._key(x, f, nf)
        return xb, fb

    # -- closed-form deconvolutions (OlaGPU.deconv) --------------------------
    def _deconv_work(self):
        if getattr(self, '_work', None) is None:
            n = int(_lib.lib().vmbd_deconv_workspace_bytes(self._h))
            self._work = torch.empty(max(n, 16), dtype=torch.uint8, device=self.device)
        return self._work

    def _spec1(self, spec):
        if self._spec_frames(spec) != 1:
            raise ValueError('the deconvolutions take one frame')
        return spec

    def deconv_direct(self, kind, spec, y, alpha, beta, out=None):
        """'xdirect' (olaGPU.py:550-609) with the cached spectrum of an 'F'
        (kind 'F': returns the image) or 'X' instance (kind 'X': the PSFs)."""
        y, nf, bat = self._g(y)
        if bat:
            raise ValueError('the deconvolutions take one frame')
        self._spec1(spec)
        k = _lib.CACHE_PSF if kind == 'F' else _lib.CACHE_IMAGE
        shape = self.sx if kind == 'F' else (self.P,) + self.sf
        o = self._empty(shape, out)
        self._rows_key = None
        _lib.check(_lib.lib().vmbd_deconv_direct(self._h, k, _lib.ptr(spec), _lib.ptr(y), float(alpha), float(beta),
                                                 _lib.ptr(o), _lib.stream_ptr()), 'deconv_direct')
        return o

    def deconv_gdirect(self, xx, xy, yx, yy, alpha, beta, out=None):
        """'gdirect' (olaGPU.py:416-511) from the valid-mode gradients of the
        latent image (xx, xy) and of the tapered observation (yx, yy)."""
        H, W = self.sx
        s0, s1 = self.sy
        want = dict(xx=(H, W - 1), xy=(H - 1, W), yx=(s0, s1 - 1), yy=(s0 - 1, s1))
        ts = {}
        for name, t in (('xx', xx), ('xy', xy), ('yx', yx), ('yy', yy)):
            t = _as_dev(t, self.device, name)
            if tuple(t.shape) != want[name]:
                raise ValueError(f'{name} must have shape {want[name]}')
            ts[name] = t
        o = self._empty((self.P,) + self.sf, out)
        self._rows_key = None
        _lib.check(_lib.lib().vmbd_deconv_gdirect(self._h, _lib.ptr(ts['xx']), _lib.ptr(ts['xy']), _lib.ptr(ts['yx']),
                                                  _lib.ptr(ts['yy']), float(alpha), float(beta), _lib.ptr(o),
                                                  _lib.ptr(self._deconv_work()), _lib.stream_ptr()), 'deconv_gdirect')
        return o

    def sparse_setup(self, spec, y, c, alpha, gamma):
        """'sparse' set-up (olaGPU.py:655-692): (F^H y spectrum, constant-image
        factor 0.01 / max(deconvolved c, 0.01)) from the tapered observation y
        and the tapered constant image c."""
        self._spec1(spec)
        y, _, by = self._g(y)
        c, _, bc = self._g(c)
        if by or bc:
            raise ValueError('the deconvolutions take one frame')
        ftpy = torch.empty_like(spec)
        cfac = torch.empty(self.sx, dtype=torch.float32, device=self.device)
        self._rows_key = None
        _lib.check(_lib.lib().vmbd_deconv_sparse_setup(self._h, _lib.ptr(spec), _lib.ptr(y), _lib.ptr(c),
                                                       float(alpha), float(gamma), _lib.ptr(ftpy), _lib.ptr(cfac),
                                                       _lib.ptr(self._deconv_work()), _lib.stream_ptr()),
                   'sparse_setup')
        return ftpy, cfac

    def sparse_step(self, spec, ftpy, wx, wy, alpha, beta, gamma, cfac, out=None):
        """One half-quadratic x-step (olaGPU.py:700-730) from the shrunk
        gradients wx, wy of the current estimate."""
        self._spec1(spec)
        self._spec1(ftpy)
        wx = _as_dev(wx, self.device, 'wx')
        wy = _as_dev(wy, self.device, 'wy')
        if wx.dim() != 2 or wy.dim() != 2:
            raise ValueError('wx and wy must be 2-D')
        cfac = self._like(cfac, self.sx, 'cfac')
        o = self._empty(self.sx, out)
        self._rows_key = None
        _lib.check(_lib.lib().vmbd_deconv_sparse_step(
            self._h, _lib.ptr(spec), _lib.ptr(ftpy), _lib.ptr(wx), wx.shape[0], wx.shape[1], _lib.ptr(wy),
            wy.shape[0], wy.shape[1], float(alpha), float(beta), float(gamma), _lib.ptr(cfac), _lib.ptr(o),
            _lib.ptr(self._deconv_work()), _lib.stream_ptr()), 'sparse_step')
        return o

    def windows(self):
        """The (P, sw0, sw1) normalised windows as the kernels apply them."""
        ws = torch.empty((self.P,) + self.sw, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().vmbd_windows(self._h, _lib.ptr(ws), _lib.stream_ptr()), 'windows')
        return ws

    # -- cached spectra (OlaGPU semantics) ----------------------------------
    def spectra(self, kind, z, out=None):
        """kind 'F': z = PSF stack; kind 'X': z = image.  Returns a complex64
        tensor of shape ([B,] P, Tw/2+1, Th) (written into `out` if given)."""
        if kind == 'F':
            z, nf, bat = self._f(z)
            k = _lib.CACHE_PSF
        elif kind == 'X':
            z, nf, bat = self._x(z)
            k = _lib.CACHE_IMAGE
        else:
            raise ValueError("kind must be 'F' or 'X'")
        shape = self._lead(nf, bat) + (self.P, self.tile[1] // 2 + 1, self.tile[0])
        if out is None:
            spec = torch.empty(shape, dtype=torch.complex64, device=self.device)
        else:
            if (tuple(out.shape) != shape or out.dtype != torch.complex64 or not out.is_contiguous()
                    or out.device != self.device):
                raise ValueError(f'out must be a contiguous complex64 tensor of shape {shape}')
            spec = out
        self._rows_key = None
        _lib.check(_lib.lib().vmbd_spectra(self._h, nf, k, _lib.ptr(z), _lib.ptr(spec), _lib.stream_ptr()),
                   'spectra')
        return spec

    def _spec_frames(self, spec):
        nf = 1 if spec.dim() == 3 else int(spec.shape[0])
        if spec.dtype != torch.complex64 or not spec.is_contiguous() or spec.numel() != nf * self.spec_elems:
            raise ValueError('spectrum tensor does not match the plan')
        return nf

    def _like(self, t, shape, name):
        if t is None:
            return None
        t = _as_dev(t, self.device, name)
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f'{name} must have shape {tuple(shape)}')
        return t

    def forward_cached(self, kind, spec, u, out=None, sub=None, mul=None):
        """B u with the cached operand; with `sub`/`mul` (shaped like y) the
        result is (B u - sub) * mul, formed in the operator's final store."""
        if kind == 'F':
            u, nf, bat = self._x(u)
            k = _lib.CACHE_PSF
        else:
            u, nf, bat = self._f(u)
            k = _lib.CACHE_IMAGE
        if self._spec_frames(spec) != nf:
            raise ValueError('spectrum and operand disagree on the number of frames')
        shape = self._lead(nf, bat) + self.sy
        sub, mul = self._like(sub, shape, 'sub'), self._like(mul, shape, 'mul')
        y = self._empty(shape, out)
        self._rows_key = None
        _lib.check(_lib.lib().vmbd_ola_forward_cached_ex(self._h, nf, k, _lib.ptr(spec), _lib.ptr(u), _lib.ptr(sub),
                                                         _lib.ptr(mul), _lib.ptr(y), _lib.stream_ptr()),
                   'forward_cached')
        return y

    def adjoint_cached(self, kind, spec, g, exact=False, out=None, z=None, beta=0.0):
        """B' g with the cached operand; with `z` (shaped like the result) the
        result is B' g + beta z, formed in the operator's final store."""
        g, nf, bat = self._g(g)
        if self._spec_frames(spec) != nf:
            raise ValueError('spectrum and residual disagree on the number of frames')
        if kind == 'F':
            k = _lib.CACHE_PSF
            shape = self._lead(nf, bat) + self.sx
        else:
            k = _lib.CACHE_IMAGE
            shape = self._lead(nf, bat) + (self.P,) + self.sf
        z = self._like(z, shape, 'z')
        o = self._empty(shape, out)
        self._rows_key = None
        _lib.check(_lib.lib().vmbd_ola_adjoint_cached_ex(self._h, nf, k, _lib.ptr(spec), _lib.ptr(g), _lib.ptr(z),
                                                         float(beta), _lib.ptr(o), int(bool(exact)),
                                                         _lib.stream_ptr()), 'adjoint_cached')
        return o
