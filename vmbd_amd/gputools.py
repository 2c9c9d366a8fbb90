"""Drop-in of the reference's gputools.py wrappers on the OLA path, backed by
the HIP kernels of libvmbd_hip.so (include/vmbd_plain.h, vmbd_solver.h,
vmbd_filters.h).

Device arrays are torch tensors (float32 or complex64) on the current HIP
device; functions that accept host arrays in the reference accept numpy
arrays and move them.  Names, argument order and defaults follow
gputools.py; line numbers are cited per function.
"""
import numpy as np
import torch

from . import _lib
from .imagetools import get_window_barthann

_C = torch.complex64


def _dev(x, dtype=torch.float32):
    """Contiguous device tensor of `dtype` (numpy arrays are copied over)."""
    if isinstance(x, torch.Tensor):
        t = x
        if t.device.type != 'cuda':
            t = t.to('cuda')
    else:
        t = torch.as_tensor(np.ascontiguousarray(x), device='cuda')
    if dtype == torch.float32 and t.is_complex():
        t = t.real
    return t.to(dtype).contiguous()


def _p(t):
    return _lib.ptr(t)


def _inplace(t, name):
    """Check that an in-place operand is a contiguous float32 device tensor
    (the kernels address it as one flat array of numel() elements) and mark
    it modified (its torch version counter moves, as with a torch in-place
    op), since the kernel about to run writes it through a raw pointer."""
    if not (isinstance(t, torch.Tensor) and t.device.type == 'cuda' and t.dtype == torch.float32
            and t.is_contiguous()):
        raise TypeError(f'{name}: expected a contiguous float32 device tensor')
    _lib.wrote(t)
    return t


def _st():
    return _lib.stream_ptr()


def _pair(v):
    a = np.asarray(v).ravel()
    if a.size == 1:
        return int(a[0]), int(a[0])
    return int(a[0]), int(a[1])


def _cplx(dtype):
    if dtype not in ('real', 'complex'):
        raise ValueError("dtype must be 'real' or 'complex'")
    return dtype == 'complex'


def _pow2(sw):
    """Default tile of chop_pad_GPU (gputools.py:326-327)."""
    return tuple(int(v) for v in 2 ** np.ceil(np.log2(np.asarray(sw, dtype=np.float64).reshape(-1) * np.ones(2))))


# --------------------------------------------------------------- chop/pad
def chop_pad_GPU(x, csf, sw, nhop, sz=None, offset=(0, 0), dtype='real'):
    """gputools.py:364-401: patches of size sw at hop nhop of x (zero padded
    to nhop*csf+sw-nhop), placed at `offset` in (prod(csf), sz0, sz1) tiles."""
    return _chop(x, None, csf, sw, nhop, sz, offset, dtype)


def chop_mod_pad_GPU(x, ws_gpu, csf, sw, nhop, sz=None, offset=(0, 0), dtype='real'):
    """gputools.py:322-361: chop_pad_GPU with each patch multiplied by its
    window ws_gpu[k] (prod(csf), sw0, sw1)."""
    return _chop(x, _dev(ws_gpu), csf, sw, nhop, sz, offset, dtype)


def _chop(x, ws, csf, sw, nhop, sz, offset, dtype):
    cplx = _cplx(dtype)
    x = _dev(x)
    sw0, sw1 = _pair(sw)
    h0, h1 = _pair(nhop)
    c0, c1 = _pair(csf)
    sz0, sz1 = _pair(_pow2((sw0, sw1)) if sz is None else sz)
    o0, o1 = _pair(offset)
    if ws is not None and tuple(ws.shape) != (c0 * c1, sw0, sw1):
        raise IOError('Size missmatch between windows and patch grid')
    out = torch.empty((c0 * c1, sz0, sz1), dtype=_C if cplx else torch.float32, device=x.device)
    _lib.call('vmbd_chop_pad', _p(out), sz0, sz1, _p(x), x.shape[0], x.shape[1], c0, c1, sw0, sw1, h0, h1, o0, o1,
              _p(ws), None, int(cplx), _st())
    return out


def pad_cpu2gpu(x, sz, offset=(0, 0), dtype='real'):
    """gputools.py:1099-1137: x placed at `offset` in a zero array of size sz."""
    cplx = _cplx(dtype)
    x = _dev(x)
    m, n = _pair(sz)
    s, t = _pair(offset)
    out = torch.empty((m, n), dtype=_C if cplx else torch.float32, device=x.device)
    _lib.call('vmbd_pad', _p(out), m, n, _p(x), x.shape[0], x.shape[1], s, t, int(cplx), _st())
    return out


def pad_stack_GPU(x, sz, offset=(0, 0), dtype='real'):
    """gputools.py:1140-1181: each slice of x (K, xm, xn) placed at `offset` in
    a zero (sz0, sz1) tile."""
    cplx = _cplx(dtype)
    x = _dev(x)
    sz0, sz1 = _pair(sz)
    o0, o1 = _pair(offset)
    out = torch.empty((x.shape[0], sz0, sz1), dtype=_C if cplx else torch.float32, device=x.device)
    _lib.call('vmbd_pad_stack', _p(out), x.shape[0], sz0, sz1, _p(x), x.shape[1], x.shape[2], o0, o1, int(cplx), _st())
    return out


def crop_stack_GPU(x, sz, offset=(0, 0), dtype='real'):
    """gputools.py:599-640: crop every slice to sz at `offset` (real part for
    dtype 'real')."""
    x = x if isinstance(x, torch.Tensor) else _dev(x)
    if x.device.type != 'cuda':
        x = x.to('cuda')
    sz0, sz1 = _pair(sz)
    o0, o1 = _pair(offset)
    if x.shape[1] - sz0 < o0 or x.shape[2] - sz1 < o1:
        raise IOError('Size missmatch: Size after - size before < offset')
    if dtype == 'complex':
        return x[:, o0:o0 + sz0, o1:o1 + sz1].to(_C).contiguous()
    cplx = x.is_complex()
    src = x.to(_C if cplx else torch.float32).contiguous()
    out = torch.empty((x.shape[0], sz0, sz1), dtype=torch.float32, device=x.device)
    _lib.call('vmbd_crop_stack', _p(out), x.shape[0], sz0, sz1, _p(src), x.shape[1], x.shape[2], o0, o1, int(cplx),
              _st())
    return out


def crop_gpu2cpu(x_gpu, sz, offset=(0, 0)):
    """gputools.py:546-569: crop to sz at `offset` (real part of complex)."""
    x = x_gpu if isinstance(x_gpu, torch.Tensor) else _dev(x_gpu)
    cplx = x.is_complex()
    src = x.to(_C if cplx else torch.float32).contiguous()
    m, n = _pair(sz)
    s, t = _pair(offset)
    out = torch.empty((m, n), dtype=torch.float32, device=src.device)
    _lib.call('vmbd_crop', _p(out), m, n, _p(src), src.shape[0], src.shape[1], s, t, int(cplx), _st())
    return out


def ola_GPU_test(xs_gpu, csf, sw, nhop, offset=(0, 0)):
    """gputools.py:1068-1096: overlap-add of the (prod(csf), m, n) stack with
    patch extent sw at hop nhop; output nhop*csf + sw - nhop."""
    xs = xs_gpu if isinstance(xs_gpu, torch.Tensor) else _dev(xs_gpu)
    cplx = xs.is_complex()
    xs = xs.to(_C if cplx else torch.float32).contiguous()
    c0, c1 = _pair(csf)
    w0, w1 = _pair(sw)
    h0, h1 = _pair(nhop)
    o0, o1 = _pair(offset)
    if xs.dim() != 3 or xs.shape[0] < c0 * c1:
        raise IOError('ola_GPU_test: the stack holds fewer than prod(csf) slices')
    out = torch.empty((h0 * c0 + w0 - h0, h1 * c1 + w1 - h1), dtype=torch.float32, device=xs.device)
    _lib.call('vmbd_ola', _p(out), _p(xs), xs.shape[1], xs.shape[2], c0, c1, w0, w1, h0, h1, o0, o1, int(cplx), _st())
    return out


# ------------------------------------------------------ elementwise helpers
EW_MUL, EW_DIV, EW_SDIV, EW_SUB, EW_AXPY, EW_SCALE, EW_LT, EW_ADDS = range(8)


def ew(op, a, b=None, s=0.0, out=None):
    """Elementwise float32 op on the device (vmbd_ew): a*b, a/b, s/a, a-b,
    a+s*b, s*a, (a<s), a+s."""
    a = _dev(a)
    b = _dev(b) if b is not None else None
    if b is not None and b.shape != a.shape:
        raise ValueError('operand shapes differ')
    o = torch.empty_like(a) if out is None else out
    _lib.call('vmbd_ew', _p(o), _p(a), _p(b), float(s), a.numel(), int(op), _st())
    return o


def cmul(a, b, mode=0, scale=1.0):
    """Complex product a*b (mode 0), a*conj(b) (1), conj(a)*b (2); b is
    broadcast over leading dimensions when it has fewer elements."""
    a = _cdev(a)
    b = _cdev(b)
    if a.numel() % b.numel():
        raise ValueError('b does not broadcast over a')
    out = torch.empty_like(a)
    _lib.call('vmbd_cmul', _p(out), _p(a), _p(b), a.numel(), b.numel(), int(mode), float(scale), _st())
    return out


def dot(a, b):
    """sum a*b as a Python float (device reduction, fixed order)."""
    a = _dev(a)
    b = _dev(b)
    n = a.numel()
    if b.numel() != n:
        raise ValueError('dot: operand sizes differ')
    part = torch.empty(3 * _lib.lib().vmbd_reduce_blocks(n), dtype=torch.float64, device=a.device)
    out = torch.empty(1, dtype=torch.float64, device=a.device)
    _lib.call('vmbd_dot', _p(a), _p(b), n, _p(part), _p(out), _st())
    return float(out.item())


# ------------------------------------------------------ Fourier divisions
def _cdev(t):
    t = t if isinstance(t, torch.Tensor) else _dev(t, _C)
    if t.device.type != 'cuda':
        t = t.to('cuda')
    return t.to(_C).contiguous()


def _check_spectra(stacks, tiles, per):
    """Per-patch spectra share one element count; broadcast operands hold one
    tile of `per` elements."""
    n = stacks[0].numel()
    if n % per or any(s.numel() != n for s in stacks):
        raise ValueError('spectrum stacks differ in size')
    if any(t.numel() != per for t in tiles):
        raise ValueError('broadcast spectra must hold exactly one tile')


def comp_ola_deconv(fs_gpu, ys_gpu, L_gpu, alpha, beta):
    """gputools.py:498-519 (gputools.cu:316-341)."""
    f, y, L = _cdev(fs_gpu), _cdev(ys_gpu), _cdev(L_gpu)
    per = f.shape[-2] * f.shape[-1]
    _check_spectra((f, y), (L,), per)
    out = torch.empty_like(f)
    _lib.call('vmbd_fourier_div', _p(out), _p(f), _p(y), _p(L), f.numel(), per, float(alpha), float(beta), _st())
    return out


def comp_ola_gdeconv(xx_gpu, xy_gpu, yx_gpu, yy_gpu, L_gpu, alpha, beta):
    """gputools.py:522-543 (gputools.cu:343-374)."""
    xx, xy, yx, yy, L = (_cdev(v) for v in (xx_gpu, xy_gpu, yx_gpu, yy_gpu, L_gpu))
    per = xx.shape[-2] * xx.shape[-1]
    _check_spectra((xx, xy, yx, yy), (L,), per)
    out = torch.empty_like(xx)
    _lib.call('vmbd_fourier_div_g', _p(out), _p(xx), _p(xy), _p(yx), _p(yy), _p(L), xx.numel(), per, float(alpha),
              float(beta), _st())
    return out


def comp_ola_sdeconv(gx_gpu, gy_gpu, xx_gpu, xy_gpu, Ftpy_gpu, f_gpu, L_gpu, alpha, beta, gamma=0):
    """gputools.py:572-596 (gputools.cu:376-406)."""
    gx, gy, xx, xy, fy, f, L = (_cdev(v) for v in (gx_gpu, gy_gpu, xx_gpu, xy_gpu, Ftpy_gpu, f_gpu, L_gpu))
    per = xx.shape[-2] * xx.shape[-1]
    _check_spectra((xx, xy, fy, f), (gx, gy, L), per)
    out = torch.empty_like(xx)
    _lib.call('vmbd_fourier_div_s', _p(out), _p(gx), _p(gy), _p(xx), _p(xy), _p(fy), _p(f), _p(L), xx.numel(), per,
              float(alpha), float(beta), float(gamma), _st())
    return out


# ------------------------------------------------------ clipping, normalise
def cliplower_GPU(x_gpu, lb):
    """gputools.py:478-485, in place."""
    _inplace(x_gpu, 'cliplower_GPU')
    _lib.call('vmbd_clip', _p(x_gpu), x_gpu.numel(), float(lb), 0.0, 1, _st())


def clipupper_GPU(x_gpu, ub):
    """gputools.py:488-495, in place."""
    _inplace(x_gpu, 'clipupper_GPU')
    _lib.call('vmbd_clip', _p(x_gpu), x_gpu.numel(), 0.0, float(ub), 2, _st())


def clip_GPU(x_gpu, lb, ub):
    """gputools.py:472-475, in place."""
    _inplace(x_gpu, 'clip_GPU')
    _lib.call('vmbd_clip', _p(x_gpu), x_gpu.numel(), float(lb), float(ub), 3, _st())


def normalize(f_gpu):
    """gputools.py:1025-1044: each PSF of a 3-D stack divided by its sum when
    the sum exceeds 1e-6; a 2-D array divided by its sum.  Device input:
    normalised on the device, a new tensor is returned; numpy input is
    normalised in place on the host, as in the reference."""
    if not isinstance(f_gpu, torch.Tensor):
        f = f_gpu
        if f.ndim == 3:
            for i in range(f.shape[0]):
                nv = f[i].sum()
                if nv > 1e-6:
                    f[i] /= nv
        else:
            f /= f.sum()
        return f
    f = _dev(f_gpu).clone()
    if f.dim() == 3:
        _lib.call('vmbd_psf_normalize', _p(f), f.shape[0], f.shape[1] * f.shape[2], 1e-6, _st())
    else:
        _lib.call('vmbd_psf_normalize', _p(f), 1, f.numel(), -float('inf'), _st())
    return f


# ----------------------------------------------------------------- filters
def impad_gpu(y_gpu, sf):
    """gputools.py:884-907: replicate padding by sf, floor(sf/2) top/left."""
    y = _dev(y_gpu)
    p0, p1 = _pair(sf)
    out = torch.empty((y.shape[0] + p0, y.shape[1] + p1), dtype=torch.float32, device=y.device)
    _lib.call('vmbd_impad', _p(out), _p(y), y.shape[0], y.shape[1], p0, p1, _st())
    return out


def edgetaper_windows(shape, sf, win='barthann'):
    """Host tables of edgetaper_gpu (gputools.py:651-659): sf made odd, 1-D
    windows of length sf, first floor(sf/2) samples / (max wx * max wy)."""
    if win != 'barthann':
        raise NotImplementedError('only the barthann window is implemented')
    sf = np.asarray(_pair(sf))
    sf = sf + (1 - np.mod(sf, 2))
    wx = get_window_barthann(int(sf[1]))
    wy = get_window_barthann(int(sf[0]))
    maxw = wx.max() * wy.max()
    hsf = np.floor(sf / 2).astype(int)
    return (wx[:hsf[1]] / maxw).astype(np.float32), (wy[:hsf[0]] / maxw).astype(np.float32)


def edgetaper_gpu(y_gpu, sf, win='barthann'):
    """gputools.py:643-673: in-place border taper of width floor(sf/2)."""
    _inplace(y_gpu, 'edgetaper_gpu')
    wx, wy = edgetaper_windows(y_gpu.shape, sf, win)
    dwx = torch.as_tensor(wx, device=y_gpu.device) if wx.size else None
    dwy = torch.as_tensor(wy, device=y_gpu.device) if wy.size else None
    _lib.call('vmbd_edgetaper', _p(y_gpu), y_gpu.shape[0], y_gpu.shape[1], _p(dwx), int(wx.size), _p(dwy),
              int(wy.size), _st())


def gradient_gpu(y_gpu, mode='valid'):
    """gputools.py:854-881: (gx, gy) forward differences / sqrt(2)."""
    y = _dev(y_gpu)
    H, W = y.shape
    if mode == 'valid':
        gx = torch.empty((H, W - 1), dtype=torch.float32, device=y.device)
        gy = torch.empty((H - 1, W), dtype=torch.float32, device=y.device)
    elif mode == 'same':
        gx = torch.empty((H, W), dtype=torch.float32, device=y.device)
        gy = torch.empty((H, W), dtype=torch.float32, device=y.device)
    else:
        raise NotImplementedError('Not a valid mode!')
    _lib.call('vmbd_gradient', _p(gx), _p(gy), _p(y), H, W, int(mode == 'same'), _st())
    return gx, gy


def laplace_gpu(y_gpu, mode='valid'):
    """gputools.py:910-933."""
    y = _dev(y_gpu)
    H, W = y.shape
    if mode == 'valid':
        out = torch.empty((H - 2, W - 2), dtype=torch.float32, device=y.device)
    elif mode == 'same':
        out = torch.empty((H, W), dtype=torch.float32, device=y.device)
    else:
        raise NotImplementedError('Not a valid mode!')
    _lib.call('vmbd_laplace', _p(out), _p(y), 1, H, W, int(mode == 'same'), 0, _st())
    return out


def laplace_stack_gpu(y_gpu, mode='valid'):
    """gputools.py:936-958 (kernel laplace_stack_same, kernel.cu:335-375: the
    'same' Laplacian of every slice, first and last slice zero; `mode` is
    ignored as in the reference)."""
    y = _dev(y_gpu)
    K, H, W = y.shape
    out = torch.empty_like(y)
    _lib.call('vmbd_laplace', _p(out), _p(y), K, H, W, 1, 1, _st())
    return out


def laplace3d_gpu(y_gpu):
    """gputools.py:961-981."""
    y = _dev(y_gpu)
    D, H, W = y.shape
    out = torch.empty_like(y)
    _lib.call('vmbd_laplace3d', _p(out), _p(y), D, H, W, _st())
    return out


def modify_sparse23_gpu(y_gpu, beta):
    """gputools.py:984-1001, in place."""
    _inplace(y_gpu, 'modify_sparse23_gpu')
    _lib.call('vmbd_shrink_alpha23', _p(y_gpu), y_gpu.numel(), float(beta), _st())


def modify_sparse_gpu(y_gpu, beta, alpha=2. / 3.):
    """gputools.py:1004-1022, in place."""
    _inplace(y_gpu, 'modify_sparse_gpu')
    _lib.call('vmbd_shrink_alpha', _p(y_gpu), y_gpu.numel(), float(beta), float(alpha), _st())


def shock_filter_gpu(y_gpu, iter=3, dt=0.1, h=1):
    """gputools.py:1282-1314: `iter` shock steps, in place (each step reads
    the previous image and writes a new one)."""
    _inplace(y_gpu, 'shock_filter_gpu')
    H, W = y_gpu.shape
    a = y_gpu
    b = torch.empty_like(y_gpu)
    for _ in range(int(iter)):
        _lib.call('vmbd_shock_step', _p(b), _p(a), H, W, float(dt), float(h), _st())
        a, b = b, a
    if a is not y_gpu:
        y_gpu.copy_(a)


def bfilter_gpu(y_gpu, w=5, sigma_d=3, sigma_r=0.1):
    """gputools.py:251-285: bilateral filter, in place."""
    _inplace(y_gpu, 'bfilter_gpu')
    H, W = y_gpu.shape
    out = torch.empty_like(y_gpu)
    _lib.call('vmbd_bilateral', _p(out), _p(y_gpu), H, W, int(w), float(sigma_d), float(sigma_r), _st())
    y_gpu.copy_(out)


def resize_gpu(y_gpu, out_shape):
    """gputools.py:1222-1266: bilinear resampling to out_shape."""
    y = _dev(y_gpu)
    oh, ow = _pair(out_shape)
    out = torch.empty((oh, ow), dtype=torch.float32, device=y.device)
    _lib.call('vmbd_resize', _p(out), oh, ow, _p(y), 1, y.shape[0], y.shape[1], _st())
    return out


def resize_stack_gpu(fs_gpu, out_shape):
    """gputools.py:1205-1219: every slice resampled (one launch)."""
    f = _dev(fs_gpu)
    oh, ow = _pair(out_shape)
    out = torch.empty((f.shape[0], oh, ow), dtype=torch.float32, device=f.device)
    _lib.call('vmbd_resize', _p(out), oh, ow, _p(f), f.shape[0], f.shape[1], f.shape[2], _st())
    return out
