"""Drop-in of the reference's olaGPU.OlaGPU (olaGPU.py:11-825) on the
MI355X kernels.

    Z = OlaGPU(z, sx, mode, winaux)      # z: PSF stack (P,sf0,sf1) -> 'F',
                                         #    image (H,W) with sx = PSF size -> 'X'
    y = Z.cnv(u)                         # F x  or  X f
    u = Z.cnvtp(y)                       # F' y or  X' y
    u = Z.deconv(y, mode=...)            # 'lbfgsb', 'gdirect', 'direct', 'xdirect', 'sparse'

Differences from the reference, all deliberate:
  * device arrays are torch tensors; numpy input to ``cnv`` gives numpy
    output, as in the reference (olaGPU.py:254-257);
  * the cached spectrum (``fft_gpu``) is the half spectrum of the real
    patches in the kernels' layout (P, Tw/2+1, Th) complex64 (the real row
    bins 0 and Tw/2 packed into bin 0, bin Tw/2 zero);
  * the closed-form deconvolutions ('xdirect', 'direct', 'gdirect',
    'sparse') form their Fourier divisions inside the operator's column pass
    on those half spectra (three launches per division); ``fused=False``
    runs the reference's sequence instead (chop, pad, full (P, Th, Tw)
    spectra built on first use by ``full_spectrum``, FFT, divide, IFFT, crop,
    OLA), one plain kernel per step;
  * ``cnvtp`` of an 'F' instance overlap-adds with the patch extent sw (the
    reference passes sw+sf-1 and reads past every patch, olaGPU.py:307-309);
    ``exact=True`` gives the true adjoint instead of the patch-local transpose;
  * inverse FFTs run the inverse transform directly instead of
    conj(FFT(conj(.))) (olaGPU.py:225-231), which is the same map.
"""
import numpy as np
import torch

from . import _lib
from . import gputools
from .imagetools import fft_tile
from .ola import OlaPlan


def _is_host(a):
    return isinstance(a, np.ndarray)


def _spectral_energy(gx, gy):
    """|Gx|^2 + |Gy|^2 as a complex tile (imaginary part 0)."""
    return gputools.cmul(gx, gx, mode=1) + gputools.cmul(gy, gy, mode=1)


class OlaGPU:
    def __init__(self, f, sx, mode, winaux, exact=False, max_frames=1, fused=True):
        sf = np.array(f.shape)[-2:]
        sx = np.array(sx)
        sw = np.asarray(winaux.sw)
        csf = winaux.csf
        self.exact = bool(exact)
        # closed-form deconvolutions: fused into the operator's passes, or the
        # reference's sequence of plain kernels over full complex spectra
        self.fused = bool(fused)
        if len(f.shape) == 3 and np.all(sx > sf):                     # olaGPU.py:104-111
            self.f = f
            self.x = []
            self.__id__ = 'F'
            if int(np.prod(f.shape[0])) != int(np.prod(csf)):
                raise IOError('Size missmatch between winaux and PSF size!')
        elif len(f.shape) == 2 and np.all(sx < sf):                   # olaGPU.py:113-118
            self.f = []
            self.x = f
            self.__id__ = 'X'
            sf = sx
            sx = np.array(self.x.shape)
        else:                                                         # olaGPU.py:120-121
            raise IOError('Size missmatch')
        if np.any(np.asarray(winaux.sx) != sx):                       # olaGPU.py:124-125
            raise IOError('Size missmatch between winaux and image size!')
        if mode == 'valid':                                           # olaGPU.py:127-136
            sy = sx - sf + 1
        elif mode in ('same', 'circ'):
            sy = sx.copy()
        elif mode == 'full':
            sy = sx + sf - 1
        else:
            raise NotImplementedError('Not a valid mode!')
        self.sfft = sw + sf - 1
        self.sfft_gpu = fft_tile(sw, sf)
        self.plan = OlaPlan(tuple(sx), tuple(sf), winaux, mode='same' if mode == 'circ' else mode,
                            tile=self.sfft_gpu, max_frames=max_frames)
        if self.__id__ == 'F':
            self.fft_gpu = self.plan.spectra('F', gputools._dev(self.f))
            self.sz = sx
        else:
            self.fft_gpu = self.plan.spectra('X', gputools._dev(self.x))
            self.sz = sf
        self.winaux = winaux
        self.sf = sf
        self.sx = sx
        self.sy = sy
        self.mode = mode
        self._full = None
        self._tw = None

    # ------------------------------------------------------------ helpers
    @property
    def _kind(self):
        return self.__id__

    def _twiddles(self):
        if self._tw is None:
            self._tw = torch.empty(2 * 1024, dtype=torch.float32, device=self.plan.device)
            _lib.call('vmbd_twiddles', _lib.ptr(self._tw), _lib.stream_ptr())
        return self._tw

    def full_spectrum(self):
        """Full (P, Th, Tw) complex64 spectrum of the cached operand, as the
        reference caches it (olaGPU.py:142-157)."""
        if self._full is None:
            wa = self.winaux
            if self.__id__ == 'F':
                z = gputools.pad_stack_GPU(self.f, self.sfft_gpu, dtype='complex')
            else:
                z = gputools.chop_mod_pad_GPU(self.x, wa.ws_gpu, wa.csf, wa.sw, wa.nhop, sz=self.sfft_gpu,
                                              dtype='complex')
            self.fft(z, z.shape[0])
            self._full = z
        return self._full

    def fft(self, y_gpu, batch_size):
        """olaGPU.py:813-824: in-place forward 2-D FFT of batch_size tiles."""
        T = y_gpu.shape[-2:]
        _lib.call('vmbd_fft2', _lib.ptr(y_gpu), int(batch_size), int(T[0]), int(T[1]), -1, 1.0,
                  _lib.ptr(self._twiddles()), _lib.stream_ptr())

    def ifft(self, y_gpu, batch_size):
        """Scaled inverse of ``fft`` (conj(fft(conj(.)))/prod(T) in the
        reference, olaGPU.py:229-231)."""
        T = y_gpu.shape[-2:]
        _lib.call('vmbd_fft2', _lib.ptr(y_gpu), int(batch_size), int(T[0]), int(T[1]), 1,
                  1.0 / float(T[0] * T[1]), _lib.ptr(self._twiddles()), _lib.stream_ptr())

    # ------------------------------------------------------------ operator
    def cnv(self, u):
        """olaGPU.py:170-257: F x (F instance) or X f (X instance)."""
        host = _is_host(u)
        if len(u.shape) == 3 and self.__id__ == 'X':
            if int(np.prod(u.shape[0])) != int(np.prod(self.winaux.csf)):
                raise IOError('Size missmatch between winaux and PSF size!')
        elif len(u.shape) == 2 and self.__id__ == 'F':
            if np.sum(np.array(u.shape) != np.asarray(self.winaux.sx)) > 0:
                raise IOError('Size missmatch between winaux and image size!')
        else:
            raise IOError('Size missmatch')
        if self.mode == 'circ':
            raise NotImplementedError('Not a valid mode!')
        y = self.plan.forward_cached(self.__id__, self.fft_gpu, gputools._dev(u))
        return y.cpu().numpy() if host else y

    def cnvtp(self, y, exact=None):
        """olaGPU.py:261-311: X' y (PSF shaped) or F' y (image shaped).  The
        reference stops at valid/same; 'full' extends the same rule (residual
        patches at tile offset 0)."""
        if self.mode not in ('valid', 'same', 'full'):
            raise NotImplementedError('Not a valid mode!')
        ex = self.exact if exact is None else bool(exact)
        return self.plan.adjoint_cached(self.__id__, self.fft_gpu, gputools._dev(y), exact=ex)

    # -------------------------------------------------------- deconvolution
    def deconv(self, y, z0=None, mode='lbfgsb', maxfun=100, alpha=0., beta=0.01, verbose=10, m=5, edgetapering=1,
               factor=3, gamma=1e-4):
        """olaGPU.py:315-736."""
        if not np.all(np.array(y.shape) == self.sy):
            raise IOError('Sizes incompatible. Expected blurred image!')
        y_gpu = gputools._dev(y).clone()
        if mode == 'lbfgsb':
            return self._deconv_lbfgsb(y_gpu, z0, maxfun, alpha, beta, verbose, m)
        if mode == 'gdirect':
            return self._deconv_gdirect(y_gpu, alpha, beta, edgetapering)
        if mode == 'direct':
            return self._deconv_direct(y_gpu, alpha, beta, edgetapering, factor)
        if mode == 'xdirect':
            return self._deconv_xdirect(y, y_gpu, alpha, beta, edgetapering)
        if mode == 'sparse':
            return self._deconv_sparse(y_gpu, alpha, gamma, edgetapering)
        raise NotImplementedError('Not a valid deconv mode!')

    def _offset(self):
        if self.mode == 'valid':
            return tuple(int(v) for v in self.sf - 1)
        if self.mode == 'same':
            return tuple(int(v) for v in np.floor(self.sf / 2))
        raise NotImplementedError('Not a valid mode!')

    def _filter_spectrum(self, taps):
        """|G|^2 of a small difference filter placed at the tile origin."""
        g = gputools.pad_cpu2gpu(np.asarray(taps, dtype=np.float32), self.sfft_gpu, dtype='complex')
        self.fft(g, 1)
        return g

    def _deconv_lbfgsb(self, y_gpu, z0, maxfun, alpha, beta, verbose, m):
        """olaGPU.py:381-412: scipy L-BFGS-B on the host, objective and
        gradient on the device."""
        from scipy.optimize import fmin_l_bfgs_b
        self.res_gpu = torch.empty_like(y_gpu)
        if self.__id__ == 'X':
            sz = (int(np.prod(self.winaux.csf)), int(self.sz[0]), int(self.sz[1]))
        else:
            sz = tuple(int(v) for v in self.sz)
        lf = int(np.prod(sz))
        if z0 is None:
            z0 = self.cnvtp(y_gpu).cpu().numpy().astype(np.float64).flatten()
        else:
            z0 = np.asarray(z0, dtype=np.float64).flatten()
        zhat = fmin_l_bfgs_b(func=self.cnvinv_objfun, x0=z0, fprime=self.cnvinv_gradfun,
                             args=[sz, y_gpu, alpha, beta], factr=10., pgtol=10e-15, maxfun=maxfun,
                             bounds=[(0., np.inf)] * lf, m=m)
        return np.reshape(zhat[0], sz), zhat[1], zhat[2]

    def cnvinv_objfun(self, z, sz, y_gpu, alpha=0., beta=0.):
        """olaGPU.py:756-786: 1/2 |y - Z z|^2 (+ 1/2 beta |z|^2)."""
        z_gpu = gputools._dev(np.reshape(z, sz).astype(np.float32))
        self.res_gpu = gputools.ew(gputools.EW_SUB, y_gpu, self.cnv(z_gpu))
        obj = 0.5 * gputools.dot(self.res_gpu, self.res_gpu)
        if alpha > 0:
            if self.__id__ == 'X':
                # the reference calls an undefined module here (olaGPU.py:775)
                raise NameError("name 'shock' is not defined")
            self.lz_gpu = gputools.laplace_gpu(z_gpu, mode='same')
            obj += 0.5 * alpha * gputools.dot(z_gpu, self.lz_gpu)
        if beta > 0:
            obj += 0.5 * beta * gputools.dot(z_gpu, z_gpu)
        return obj

    def cnvinv_gradfun(self, z, sz, y_gpu, alpha=0., beta=0.):
        """olaGPU.py:788-811: -(Z' res + alpha L z + beta z), the sign the
        reference returns."""
        z_gpu = gputools._dev(np.reshape(z, sz).astype(np.float32))
        grad = self.cnvtp(self.res_gpu)
        if alpha > 0:
            grad = gputools.ew(gputools.EW_AXPY, grad, self.lz_gpu, alpha)
        if beta > 0:
            grad = gputools.ew(gputools.EW_AXPY, grad, z_gpu, beta)
        return -grad.cpu().numpy().astype(np.float64).flatten()

    def _deconv_gdirect(self, y_gpu, alpha, beta, edgetapering):
        """olaGPU.py:416-511: PSFs from gradient images by one Fourier division."""
        if self.__id__ != 'X':
            raise Exception('Use direct mode for image estimation!')
        self._offset()
        if self.fused:
            if edgetapering == 1:
                gputools.edgetaper_gpu(y_gpu, 2 * self.sf, 'barthann')
            xx, xy = gputools.gradient_gpu(gputools._dev(self.x))
            yx, yy = gputools.gradient_gpu(y_gpu)
            return self.plan.deconv_gdirect(xx, xy, yx, yy, alpha, beta)
        full = self.full_spectrum()
        if alpha > 0.:
            gx = self._filter_spectrum([[-1, 1], [-1, 1], [-1, 1]])
            gy = self._filter_spectrum([[-1, -1, -1], [1, 1, 1]])
            L = _spectral_energy(gx, gy)
        else:
            L = torch.zeros(full.shape[-2:], dtype=torch.complex64, device=full.device)
        if edgetapering == 1:
            gputools.edgetaper_gpu(y_gpu, 2 * self.sf, 'barthann')
        x_gpu = gputools._dev(self.x)
        xx, xy = gputools.gradient_gpu(x_gpu)
        yx, yy = gputools.gradient_gpu(y_gpu)
        wa = self.winaux
        off = self._offset()
        yx = gputools.chop_pad_GPU(yx, wa.csf, wa.sw, wa.nhop, self.sfft_gpu, off, 'complex')
        yy = gputools.chop_pad_GPU(yy, wa.csf, wa.sw, wa.nhop, self.sfft_gpu, off, 'complex')
        xx = gputools.chop_pad_GPU(xx, wa.csf, wa.sw, wa.nhop, self.sfft_gpu, dtype='complex')
        xy = gputools.chop_pad_GPU(xy, wa.csf, wa.sw, wa.nhop, self.sfft_gpu, dtype='complex')
        P = full.shape[0]
        for t in (yx, yy, xx, xy):
            self.fft(t, P)
        z = gputools.comp_ola_gdeconv(xx, xy, yx, yy, L, alpha, beta)
        self.ifft(z, P)
        return gputools.crop_stack_GPU(z, self.sf)

    def _deconv_direct(self, y_gpu, alpha, beta, edgetapering, factor):
        """olaGPU.py:515-546: xdirect of y divided by xdirect of a constant
        image, border re-padded."""
        const = torch.ones_like(y_gpu)
        yd = self.deconv(y_gpu, mode='xdirect', alpha=alpha, beta=beta, edgetapering=edgetapering)
        gputools.cliplower_GPU(yd, 0)
        if edgetapering == 1:
            gputools.edgetaper_gpu(const, 2 * self.sf, 'barthann')
        const = self.deconv(const, mode='xdirect', alpha=alpha, beta=beta, edgetapering=edgetapering)
        gputools.edgetaper_gpu(const, 2 * self.sf, 'barthann')
        gputools.clip_GPU(const, 0.01, 10.)
        yd = gputools.ew(gputools.EW_DIV, yd, const)
        sz = np.array(yd.shape)
        yd = gputools.crop_gpu2cpu(yd, sz - factor * self.sf - 1, offset=np.floor((factor * self.sf - 1) / 2.))
        return gputools.impad_gpu(yd, tuple(sz - np.array(yd.shape)))

    def _deconv_xdirect(self, y, y_gpu, alpha, beta, edgetapering):
        """olaGPU.py:550-609: one Fourier division per patch.  As in the
        reference, the tapered copy is not the one that gets chopped
        (olaGPU.py:566 vs 577) and the windows sit at tile offset 0 while the
        residual patches sit at the mode offset (olaGPU.py:577-581)."""
        off = self._offset()
        if self.fused:
            # the tapered copy is never chopped: nothing to taper
            return self.plan.deconv_direct(self.__id__, self.fft_gpu, gputools._dev(y), alpha, beta)
        full = self.full_spectrum()
        if alpha > 0.:
            gx = self._filter_spectrum([[-1, 1]])
            gy = self._filter_spectrum([[-1], [1]])
            L = _spectral_energy(gx, gy)
        else:
            L = torch.zeros(full.shape[-2:], dtype=torch.complex64, device=full.device)
        if edgetapering == 1:
            gputools.edgetaper_gpu(y_gpu, 3 * self.sf, 'barthann')
        wa = self.winaux
        yp = gputools.chop_pad_GPU(y, wa.csf, wa.sw, wa.nhop, self.sfft_gpu, off, 'complex')
        ws = gputools.pad_stack_GPU(wa.ws_gpu, self.sfft_gpu, dtype='complex')
        yp = gputools.cmul(ws, yp)
        P = full.shape[0]
        self.fft(yp, P)
        z = gputools.comp_ola_deconv(full, yp, L, alpha, beta)
        self.ifft(z, P)
        if self.__id__ == 'X':
            return gputools.crop_stack_GPU(z, self.sf)
        zs = gputools.crop_stack_GPU(z, wa.sw)
        zc = gputools.ola_GPU_test(zs, wa.csf, wa.sw, wa.nhop)
        return gputools.crop_gpu2cpu(zc, self.sx)

    def _deconv_sparse(self, y_gpu, alpha, gamma, edgetapering):
        """olaGPU.py:613-733: hyper-Laplacian (|grad x|^(2/3)) deconvolution
        by half-quadratic splitting (Krishnan & Fergus, NIPS 2009)."""
        if self.fused:
            return self._deconv_sparse_fused(y_gpu, alpha, gamma, edgetapering)
        full = self.full_spectrum()
        r = np.sqrt(2.) / 2.
        gx = self._filter_spectrum(r * np.array([[-1, 1]]))
        gy = self._filter_spectrum(r * np.array([[-1], [1]]))
        L = _spectral_energy(gx, gy)
        const = torch.ones_like(y_gpu)
        if edgetapering == 1:
            gputools.edgetaper_gpu(y_gpu, 2 * self.sf, 'barthann')
            gputools.edgetaper_gpu(const, 2 * self.sf, 'barthann')
        beta = 1.
        beta_rate = 2. * np.sqrt(2.)
        beta_max = 2. ** 8
        x_gpu = y_gpu.clone()
        off = self._offset()
        wa = self.winaux
        P = full.shape[0]
        yp = gputools.chop_pad_GPU(y_gpu, wa.csf, wa.sw, wa.nhop, self.sfft_gpu, off, 'complex')
        cp = gputools.chop_pad_GPU(const, wa.csf, wa.sw, wa.nhop, self.sfft_gpu, off, 'complex')
        ws = gputools.pad_stack_GPU(wa.ws_gpu, self.sfft_gpu, off, dtype='complex')
        yp = gputools.cmul(yp, ws)
        cp = gputools.cmul(cp, ws)
        self.fft(cp, P)
        cp = gputools.comp_ola_deconv(full, cp, L, alpha, gamma)
        self.ifft(cp, P)
        cs = gputools.ew(gputools.EW_MUL, gputools.crop_stack_GPU(cp, wa.sw), wa.ws_gpu)
        const = gputools.crop_gpu2cpu(gputools.ola_GPU_test(cs, wa.csf, wa.sw, wa.nhop), self.sx)
        gputools.cliplower_GPU(const, 0.01)
        const = gputools.ew(gputools.EW_SDIV, const, s=0.01)
        self.fft(yp, P)
        yp = gputools.cmul(yp, full, mode=1)
        while beta < beta_max:
            xx, xy = gputools.gradient_gpu(x_gpu)
            gputools.modify_sparse23_gpu(xx, beta)
            gputools.modify_sparse23_gpu(xy, beta)
            xx = gputools.chop_pad_GPU(xx, wa.csf, wa.sw, wa.nhop, self.sfft_gpu, dtype='complex')
            xy = gputools.chop_pad_GPU(xy, wa.csf, wa.sw, wa.nhop, self.sfft_gpu, dtype='complex')
            self.fft(xx, P)
            self.fft(xy, P)
            z = gputools.comp_ola_sdeconv(gx, gy, xx, xy, yp, full, L, alpha, beta, gamma)
            self.ifft(z, P)
            zs = gputools.ew(gputools.EW_MUL, gputools.crop_stack_GPU(z, wa.sw), wa.ws_gpu)
            x_gpu = gputools.crop_gpu2cpu(gputools.ola_GPU_test(zs, wa.csf, wa.sw, wa.nhop), self.sx)
            x_gpu = gputools.ew(gputools.EW_MUL, x_gpu, const)
            gputools.cliplower_GPU(x_gpu, 0.)
            beta *= beta_rate
        return x_gpu

    def _deconv_sparse_fused(self, y_gpu, alpha, gamma, edgetapering):
        """The same splitting on the fused passes: one set-up (F^H y and the
        constant-image factor) and, per beta, the shrunk gradients of the
        estimate through one row pass, one division and one overlap-add."""
        if self.__id__ != 'F':
            raise NotImplementedError('sparse deconvolution estimates the image (F instance)')
        self._offset()
        const = torch.ones_like(y_gpu)
        if edgetapering == 1:
            gputools.edgetaper_gpu(y_gpu, 2 * self.sf, 'barthann')
            gputools.edgetaper_gpu(const, 2 * self.sf, 'barthann')
        ftpy, cfac = self.plan.sparse_setup(self.fft_gpu, y_gpu, const, alpha, gamma)
        beta, beta_rate, beta_max = 1., 2. * np.sqrt(2.), 2. ** 8
        x_gpu = y_gpu.clone()
        out = torch.empty(tuple(int(v) for v in self.sx), dtype=torch.float32, device=y_gpu.device)
        while beta < beta_max:
            xx, xy = gputools.gradient_gpu(x_gpu)
            gputools.modify_sparse23_gpu(xx, beta)
            gputools.modify_sparse23_gpu(xy, beta)
            x_gpu = self.plan.sparse_step(self.fft_gpu, ftpy, xx, xy, alpha, beta, gamma, cfac, out=out)
            beta *= beta_rate
        return x_gpu

    def deconv_rgb(self, y, z0=None, mode='lbfgsb', maxfun=100, iprint=10, m=5, alpha=0., beta=0.01,
                   edgetapering=1):
        """olaGPU.py:739-753, per colour channel (keyword arguments: the
        reference passes them positionally in the wrong order, :751-752)."""
        out = np.empty((int(self.sx[0]), int(self.sx[1]), 3), dtype=np.float32)
        for i in range(3):
            yt = np.ascontiguousarray(y[..., i].astype(np.float32))
            r = self.deconv(yt, z0=z0, mode=mode, maxfun=maxfun, alpha=alpha, beta=beta, verbose=iprint, m=m,
                            edgetapering=edgetapering)
            out[..., i] = r.cpu().numpy() if isinstance(r, torch.Tensor) else r
        return out
