"""FP64 NumPy restatement of OlaGPU.deconv's closed-form modes
(olaGPU.py:416-733) and the Fourier divisions of gputools.cu:316-406
(test infrastructure only).

Every step follows the reference order, including its quirks: xdirect tapers
a copy of y but chops the untapered input (olaGPU.py:566 vs 577) and places
the windows at tile offset 0 while the patches sit at the mode offset
(:577-581); 'direct' crops factor*sf-1 and re-pads by replication (:542-544).
"""
import numpy as np

from . import filters_oracle as FO
from . import ola_oracle as O


def fourier_div(f, y, L, alpha, beta):
    """gputools.cu:316-341: conj(f) y / (|f|^2 + alpha L + beta)."""
    return np.conj(f) * y / (np.abs(f) ** 2 + alpha * L + beta)


def fourier_div_g(xx, xy, yx, yy, L, alpha, beta):
    """gputools.cu:343-374."""
    return (np.conj(xx) * yx + np.conj(xy) * yy) / (np.abs(xx) ** 2 + np.abs(xy) ** 2 + alpha * L + beta)


def fourier_div_s(gx, gy, wx, wy, ftpy, f, L, alpha, beta, gamma):
    """gputools.cu:376-406."""
    ab = alpha * beta
    return (np.conj(gx) * wx + np.conj(gy) * wy + ftpy / ab) / (np.abs(f) ** 2 / ab + L + gamma)


def _filter_spectrum(taps, T):
    g = np.zeros(T, dtype=np.complex128)
    taps = np.asarray(taps, dtype=np.float64)
    g[:taps.shape[0], :taps.shape[1]] = taps
    return O.fft2(g)


def _offset(Z):
    if Z.mode == 'valid':
        return tuple(int(v) for v in Z.sf - 1)
    if Z.mode == 'same':
        return tuple(int(v) for v in np.floor(Z.sf / 2))
    raise NotImplementedError('Not a valid mode!')


def xdirect(Z, y, alpha, beta, edgetapering=1):
    """olaGPU.py:550-609 for an OlaOracle Z."""
    wa = Z.winaux
    T = Z.T
    if alpha > 0:
        gx = _filter_spectrum([[-1, 1]], T)
        gy = _filter_spectrum([[-1], [1]], T)
        L = np.abs(gx) ** 2 + np.abs(gy) ** 2
    else:
        L = np.zeros(T)
    off = _offset(Z)
    yp = O.chop_pad(np.asarray(y, dtype=np.float64), wa.csf, wa.sw, wa.nhop, T, off)
    ws = O.pad_stack(wa.ws, T)
    yp = O.fft2(ws * yp)
    z = np.real(O.ifft2(fourier_div(Z.spec, yp, L, alpha, beta)))
    if Z.id == 'X':
        return O.crop_stack(z, Z.sf)
    zs = O.crop_stack(z, wa.sw)
    return O.crop(O.ola(zs, wa.csf, wa.sw, wa.nhop), Z.sx)


def direct(Z, y, alpha, beta, edgetapering=1, factor=3):
    """olaGPU.py:515-546."""
    const = np.ones(np.shape(y))
    yd = np.maximum(xdirect(Z, y, alpha, beta, edgetapering), 0.0)
    if edgetapering == 1:
        const = FO.edgetaper(const, 2 * Z.sf)
    cd = xdirect(Z, const, alpha, beta, edgetapering)
    cd = FO.edgetaper(cd, 2 * Z.sf)
    cd = np.clip(cd, 0.01, 10.0)
    yd = yd / cd
    sz = np.array(yd.shape)
    yc = O.crop(yd, sz - factor * Z.sf - 1, np.floor((factor * Z.sf - 1) / 2.0).astype(int))
    return FO.impad(yc, tuple(sz - np.array(yc.shape)))


def gdirect(Z, y, alpha, beta, edgetapering=1):
    """olaGPU.py:416-511 (X instance)."""
    if Z.id != 'X':
        raise Exception('Use direct mode for image estimation!')
    wa = Z.winaux
    T = Z.T
    if alpha > 0:
        gx = _filter_spectrum([[-1, 1], [-1, 1], [-1, 1]], T)
        gy = _filter_spectrum([[-1, -1, -1], [1, 1, 1]], T)
        L = np.abs(gx) ** 2 + np.abs(gy) ** 2
    else:
        L = np.zeros(T)
    y = np.asarray(y, dtype=np.float64)
    if edgetapering == 1:
        y = FO.edgetaper(y, 2 * Z.sf)
    xx, xy = FO.gradient(Z.x)
    yx, yy = FO.gradient(y)
    off = _offset(Z)
    yx = O.fft2(O.chop_pad(yx, wa.csf, wa.sw, wa.nhop, T, off))
    yy = O.fft2(O.chop_pad(yy, wa.csf, wa.sw, wa.nhop, T, off))
    xx = O.fft2(O.chop_pad(xx, wa.csf, wa.sw, wa.nhop, T))
    xy = O.fft2(O.chop_pad(xy, wa.csf, wa.sw, wa.nhop, T))
    z = np.real(O.ifft2(fourier_div_g(xx, xy, yx, yy, L, alpha, beta)))
    return O.crop_stack(z, Z.sf)


def sparse(Z, y, alpha, gamma=1e-4, edgetapering=1):
    """olaGPU.py:613-733: half-quadratic splitting with the |.|^(2/3) w-step."""
    wa = Z.winaux
    T = Z.T
    r = np.sqrt(2.0) / 2.0
    gx = _filter_spectrum(r * np.array([[-1, 1]]), T)
    gy = _filter_spectrum(r * np.array([[-1], [1]]), T)
    L = np.abs(gx) ** 2 + np.abs(gy) ** 2
    y = np.asarray(y, dtype=np.float64)
    const = np.ones(y.shape)
    if edgetapering == 1:
        y = FO.edgetaper(y, 2 * Z.sf)
        const = FO.edgetaper(const, 2 * Z.sf)
    beta, beta_rate, beta_max = 1.0, 2.0 * np.sqrt(2.0), 2.0 ** 8
    x = y.copy()
    off = _offset(Z)
    ws = O.pad_stack(wa.ws, T, off)
    yp = O.chop_pad(y, wa.csf, wa.sw, wa.nhop, T, off) * ws
    cp = O.chop_pad(const, wa.csf, wa.sw, wa.nhop, T, off) * ws
    cp = np.real(O.ifft2(fourier_div(Z.spec, O.fft2(cp), L, alpha, gamma)))
    cs = O.crop_stack(cp, wa.sw) * wa.ws
    const = O.crop(O.ola(cs, wa.csf, wa.sw, wa.nhop), Z.sx)
    const = 0.01 / np.maximum(const, 0.01)
    ftpy = O.fft2(yp) * np.conj(Z.spec)
    while beta < beta_max:
        xx, xy = FO.gradient(x)
        xx = FO.shrink23(xx, beta)
        xy = FO.shrink23(xy, beta)
        wx = O.fft2(O.chop_pad(xx, wa.csf, wa.sw, wa.nhop, T))
        wy = O.fft2(O.chop_pad(xy, wa.csf, wa.sw, wa.nhop, T))
        z = np.real(O.ifft2(fourier_div_s(gx, gy, wx, wy, ftpy, Z.spec, L, alpha, beta, gamma)))
        zs = O.crop_stack(z, wa.sw) * wa.ws
        x = O.crop(O.ola(zs, wa.csf, wa.sw, wa.nhop), Z.sx) * const
        x = np.maximum(x, 0.0)
        beta *= beta_rate
    return x


def cnvinv_objfun(Z, z, y, alpha=0.0, beta=0.0):
    """olaGPU.py:756-786: 1/2 |y - Z z|^2 (+ 1/2 alpha z.Lz with L the
    'same' 5-point Laplacian of an image z, + 1/2 beta |z|^2).  An 'X'
    instance with alpha > 0 calls an undefined module in the reference
    (olaGPU.py:775): NameError here too."""
    z = np.asarray(z, dtype=np.float64)
    r = np.asarray(y, dtype=np.float64) - Z.cnv(z)
    obj = 0.5 * np.sum(r * r)
    if alpha > 0:
        if Z.id == 'X':
            raise NameError("name 'shock' is not defined")
        obj += 0.5 * alpha * np.sum(z * FO.laplace(z, 'same'))
    if beta > 0:
        obj += 0.5 * beta * np.sum(z * z)
    return obj


def cnvinv_gradfun(Z, z, y, alpha=0.0, beta=0.0, exact=False):
    """olaGPU.py:788-811: -(Z'(y - Z z) + alpha L z + beta z), the sign the
    reference returns: the data term enters as the true gradient, the two
    Tikhonov terms with the opposite sign.  Z' is the patch-local transpose
    (exact=False, OlaGPU's) or the true adjoint."""
    z = np.asarray(z, dtype=np.float64)
    r = np.asarray(y, dtype=np.float64) - Z.cnv(z)
    g = Z.cnvtp(r, exact=exact)
    if alpha > 0:
        if Z.id == 'X':
            raise NameError("name 'shock' is not defined")
        g = g + alpha * FO.laplace(z, 'same')
    if beta > 0:
        g = g + beta * z
    return -g
