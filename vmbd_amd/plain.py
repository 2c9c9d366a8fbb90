"""The plain per-operator GPU form of the EFF/OLA step: the reference's
call sequence (olaGPU.py:95-311) with one HIP kernel per operation (the
plain-decomposition kernels of include/vmbd_plain.h).

It exists as the comparator of the fused operator: the same launches, full
complex (T x T) spectra and HBM round trips per operation that the reference
issues, without its defects (the OLA of the x-transpose uses the patch
extent sw, not sw+sf-1, olaGPU.py:307-309; inverse FFTs are direct instead
of conj-FFT-conj, which is the same arithmetic).

    p = PlainOla(sx, sf, winaux)
    y, xbar, fbar = p.step(x, f, g)      # F ctor, X ctor, F.cnv, F.cnvtp, X.cnvtp

Valid mode; the transposes are the reference's patch-local ones,
z = IFFT(G^ . conj(C^)) with the e^{-i} forward transform used here: the
correlation z[s] = sum_b G[b] c[b - s] that makes X.cnvtp the transpose of
X.cnv on the patch (olaGPU.py:290-297 writes conj(G^) . C^ for pyfft's
transform; with an e^{-i} forward transform that product would correlate in
the opposite direction).
"""
import numpy as np
import torch

from . import _lib
from . import gputools as G
from .imagetools import fft_tile


class PlainOla:
    def __init__(self, sx, sf, winaux, device=None):
        self.sx = tuple(int(v) for v in sx)
        self.sf = tuple(int(v) for v in sf)
        self.sy = (self.sx[0] - self.sf[0] + 1, self.sx[1] - self.sf[1] + 1)
        self.wa = winaux
        self.sw = tuple(int(v) for v in np.broadcast_to(winaux.sw, (2,)))
        self.nhop = tuple(int(v) for v in np.broadcast_to(winaux.nhop, (2,)))
        self.csf = tuple(int(v) for v in winaux.csf)
        self.T = tuple(int(v) for v in fft_tile(self.sw, self.sf))
        self.ext = (self.sw[0] + self.sf[0] - 1, self.sw[1] + self.sf[1] - 1)
        self.device = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
        self.tw = torch.empty(2 * 1024, dtype=torch.float32, device=self.device)
        _lib.call('vmbd_twiddles', _lib.ptr(self.tw), _lib.stream_ptr())
        self.ws = winaux.ws_gpu

    def _fft(self, z, inverse=False):
        T = self.T
        _lib.call('vmbd_fft2', _lib.ptr(z), int(z.shape[0]), T[0], T[1], 1 if inverse else -1,
                  1.0 / (T[0] * T[1]) if inverse else 1.0, _lib.ptr(self.tw), _lib.stream_ptr())
        return z

    def psf_spectra(self, f):
        """F instance (olaGPU.py:142-146): FFT of the zero-padded PSFs."""
        return self._fft(G.pad_stack_GPU(f, self.T, dtype='complex'))

    def image_spectra(self, x):
        """X instance (olaGPU.py:148-157): FFT of the windowed patches."""
        return self._fft(G.chop_mod_pad_GPU(x, self.ws, self.csf, self.sw, self.nhop, self.T, dtype='complex'))

    def cnv_F(self, Fh, x):
        """F.cnv(x), olaGPU.py:196-244 (valid)."""
        u = G.chop_pad_GPU(x, self.csf, self.sw, self.nhop, self.T, (0, 0), 'complex')
        ws = G.pad_stack_GPU(self.ws, self.T, (0, 0), 'complex')
        u = self._fft(G.cmul(ws, u))
        z = self._fft(G.cmul(Fh, u), inverse=True)
        yo = G.ola_GPU_test(z, self.csf, self.ext, self.nhop)
        return G.crop_gpu2cpu(yo, self.sy, (self.sf[0] - 1, self.sf[1] - 1))

    def _residual_spectra(self, g):
        gp = G.chop_pad_GPU(g, self.csf, self.sw, self.nhop, self.T, (self.sf[0] - 1, self.sf[1] - 1), 'complex')
        return self._fft(gp)

    def cnvtp_F(self, Fh, g):
        """F.cnvtp(g), olaGPU.py:275-311 (valid), OLA over the patch extent sw."""
        z = self._fft(G.cmul(self._residual_spectra(g), Fh, mode=1), inverse=True)
        zs = G.crop_stack_GPU(z, self.sw)
        zs = G.ew(G.EW_MUL, zs, self.ws)
        return G.crop_gpu2cpu(G.ola_GPU_test(zs, self.csf, self.sw, self.nhop), self.sx)

    def cnvtp_X(self, Xh, g):
        """X.cnvtp(g), olaGPU.py:275-302 (valid)."""
        z = self._fft(G.cmul(self._residual_spectra(g), Xh, mode=1), inverse=True)
        return G.crop_stack_GPU(z, self.sf)

    def step(self, x, f, g):
        """Forward and both transposes of one frame, as the reference builds
        them: F instance, X instance, F.cnv(x), F.cnvtp(g), X.cnvtp(g)."""
        Fh = self.psf_spectra(f)
        Xh = self.image_spectra(x)
        return self.cnv_F(Fh, x), self.cnvtp_F(Fh, g), self.cnvtp_X(Xh, g)
