"""Operator plans over torch tensors: the EFF/OLA forward and its transposes
on the HIP kernels of libvmbd_hip.so.

    plan = OlaPlan(sx, sf, winaux, mode='valid')
    y    = plan.forward(x, f)             # B(x,f) = F x = X f
    xbar = plan.adjoint_x(g, f)           # F' g
    fbar = plan.adjoint_f(g, x)           # X' g
    xbar, fbar = plan.adjoint_xf(g, x, f)

All tensors are float32, contiguous, on the plan's device; results are
enqueued on the current torch stream.  ``exact=True`` gives the true adjoint,
``exact=False`` the reference's patch-local transpose (olaGPU.py:276-311).
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
    def __init__(self, sx, sf, winaux, mode='valid', tile=None, device=None):
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
        self.tile = tuple(int(v) for v in (tile if tile is not None else fft_tile(self.sw, sf)))
        g = _lib.VmbdGeom(sx[0], sx[1], sf[0], sf[1], self.csf[0], self.csf[1], self.sw[0], self.sw[1],
                          self.nhop[0], self.nhop[1], self.tile[0], self.tile[1], _lib.MODES[mode])
        table = np.ascontiguousarray(winaux.win_table, dtype=np.float32)
        idx = np.ascontiguousarray(winaux.win_idx, dtype=np.int32)
        L = _lib.lib()
        with torch.cuda.device(self.device):
            h = L.vmbd_plan_create(ctypes.byref(g), table.ctypes.data_as(ctypes.c_void_p), int(table.shape[0]),
                                   idx.ctypes.data_as(ctypes.c_void_p))
        if not h:
            raise _lib.VmbdError(f'vmbd_plan_create: {_lib.last_error()}')
        self._h = ctypes.c_void_p(h)
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(L.vmbd_plan_out_shape(self._h, ctypes.byref(a), ctypes.byref(b)), 'out_shape')
        self.sy = (a.value, b.value)
        self.spec_elems = int(L.vmbd_plan_spectrum_elems(self._h))
        self.workspace_bytes = int(L.vmbd_plan_workspace_bytes(self._h))

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
    def _x(self, x):
        x = _as_dev(x, self.device, 'x')
        if tuple(x.shape) != self.sx:
            raise IOError('Size missmatch between winaux and image size!')
        return x

    def _f(self, f):
        f = _as_dev(f, self.device, 'f')
        if f.dim() != 3 or f.shape[0] != self.P:
            raise IOError('Size missmatch between winaux and PSF size!')
        if tuple(f.shape[1:]) != self.sf:
            raise IOError('Size missmatch')
        return f

    def _g(self, g):
        g = _as_dev(g, self.device, 'g')
        if tuple(g.shape) != self.sy:
            raise IOError('Sizes incompatible. Expected blurred image!')
        return g

    def _empty(self, shape, out):
        if out is None:
            return torch.empty(shape, dtype=torch.float32, device=self.device)
        if tuple(out.shape) != tuple(shape) or out.dtype != torch.float32 or not out.is_contiguous():
            raise ValueError('out has the wrong shape/dtype/layout')
        return out

    # -- operators ----------------------------------------------------------
    def forward(self, x, f, out=None):
        x, f = self._x(x), self._f(f)
        y = self._empty(self.sy, out)
        _lib.check(_lib.lib().vmbd_ola_forward(self._h, _lib.ptr(x), _lib.ptr(f), _lib.ptr(y), _lib.stream_ptr()),
                   'forward')
        return y

    def adjoint_x(self, g, f, exact=True, out=None):
        g, f = self._g(g), self._f(f)
        xb = self._empty(self.sx, out)
        _lib.check(_lib.lib().vmbd_ola_adjoint_x(self._h, _lib.ptr(g), _lib.ptr(f), _lib.ptr(xb), int(bool(exact)),
                                                 _lib.stream_ptr()), 'adjoint_x')
        return xb

    def adjoint_f(self, g, x, exact=True, out=None):
        g, x = self._g(g), self._x(x)
        fb = self._empty((self.P,) + self.sf, out)
        _lib.check(_lib.lib().vmbd_ola_adjoint_f(self._h, _lib.ptr(g), _lib.ptr(x), _lib.ptr(fb), int(bool(exact)),
                                                 _lib.stream_ptr()), 'adjoint_f')
        return fb

    def adjoint_xf(self, g, x, f, exact=True, out_x=None, out_f=None):
        g, x, f = self._g(g), self._x(x), self._f(f)
        xb = self._empty(self.sx, out_x)
        fb = self._empty((self.P,) + self.sf, out_f)
        _lib.check(_lib.lib().vmbd_ola_adjoint_xf(self._h, _lib.ptr(g), _lib.ptr(x), _lib.ptr(f), _lib.ptr(xb),
                                                  _lib.ptr(fb), int(bool(exact)), _lib.stream_ptr()), 'adjoint_xf')
        return xb, fb

    # -- cached spectra (OlaGPU semantics) ----------------------------------
    def spectra(self, kind, z):
        """kind 'F': z = PSF stack; kind 'X': z = image.  Returns a complex64
        tensor of shape (P, Tw/2+1, Th)."""
        if kind == 'F':
            z = self._f(z)
            k = _lib.CACHE_PSF
        elif kind == 'X':
            z = self._x(z)
            k = _lib.CACHE_IMAGE
        else:
            raise ValueError("kind must be 'F' or 'X'")
        spec = torch.empty((self.P, self.tile[1] // 2 + 1, self.tile[0]), dtype=torch.complex64, device=self.device)
        _lib.check(_lib.lib().vmbd_spectra(self._h, k, _lib.ptr(z), _lib.ptr(spec), _lib.stream_ptr()), 'spectra')
        return spec

    def forward_cached(self, kind, spec, u, out=None):
        if kind == 'F':
            u = self._x(u)
            k = _lib.CACHE_PSF
        else:
            u = self._f(u)
            k = _lib.CACHE_IMAGE
        y = self._empty(self.sy, out)
        _lib.check(_lib.lib().vmbd_ola_forward_cached(self._h, k, _lib.ptr(spec), _lib.ptr(u), _lib.ptr(y),
                                                      _lib.stream_ptr()), 'forward_cached')
        return y

    def adjoint_cached(self, kind, spec, g, exact=False, out=None):
        g = self._g(g)
        if kind == 'F':
            k = _lib.CACHE_PSF
            o = self._empty(self.sx, out)
        else:
            k = _lib.CACHE_IMAGE
            o = self._empty((self.P,) + self.sf, out)
        _lib.check(_lib.lib().vmbd_ola_adjoint_cached(self._h, k, _lib.ptr(spec), _lib.ptr(g), _lib.ptr(o),
                                                      int(bool(exact)), _lib.stream_ptr()), 'adjoint_cached')
        return o
