"""FP64 NumPy restatement of the reference's EFF/OLA operator (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use this module, as the checker.  It follows the reference
line by line in *meaning* (index definitions, crop offsets, window formula) but
computes everything in float64 / complex128 with ``numpy.fft``.

Reference files cited below are relative to the reference repository root.

Parity pins (each checked in ``tests/test_oracle_*.py``):
  * ``barthann_periodic``/``window``  -- scipy.signal.get_window closed form.
  * ``win2winaux``                    -- partition of unity, geometry closed form.
  * ``pad/chop_pad/pad_stack/crop_stack/crop/ola`` -- brute-force loops that
    transcribe the CUDA index expressions one by one, and chop->OLA round trip.
  * ``OlaOracle.cnv``                 -- scipy.signal.convolve2d for an
    invariant PSF (partition of unity), brute-force spatial EFF sum,
    delta-PSF shift, ``Xf == Fx``, tile-size independence.
  * exact adjoints                    -- dot-product test against ``cnv`` and
    brute-force correlate2d; FP64 central finite differences.
  * reference-local adjoints (``exact=False``) -- brute-force correlate2d on the
    masked residual patch (the index formula of olaGPU.py:276-311).
"""
import numpy as np

EPS_WINDOW = 1e-16  # imagetools.py:342 floor


# --------------------------------------------------------------------------
# Windows and geometry
# --------------------------------------------------------------------------
def barthann_periodic(n):
    """scipy.signal.get_window('barthann', n): periodic (fftbins=True) form.

    Used by imagetools.py:332/335 and gputools.py:653-654.  Periodic = the
    symmetric window of length n+1 with its last sample dropped.
    """
    n = int(n)
    if n <= 1:
        return np.ones(n)
    m = n + 1
    k = np.arange(m, dtype=np.float64)
    fac = np.abs(k / (m - 1.0) - 0.5)
    w = 0.62 - 0.48 * fac + 0.38 * np.cos(2.0 * np.pi * fac)
    return w[:n]


def window(sw):
    """imagetools.py:314-344: symmetrised, min-shifted, outer product, /max,
    floored at 1e-16."""
    w1 = barthann_periodic(int(sw[0]))
    w1 = (w1 + w1[::-1]) / 2.0
    w1 = w1 - w1.min()
    w2 = barthann_periodic(int(sw[1]))
    w2 = (w2 + w2[::-1]) / 2.0
    w2 = w2 - w2.min()
    www = np.outer(w1, w2)
    www = www / www.max()
    return np.maximum(www, EPS_WINDOW)


def geometry(sx, csf, overlap):
    """imagetools.py:423-436: window size ``sw`` and hop ``nhop`` (both dims
    incremented together until ``sw*overlap`` is integral; bail back to the
    first guess after 2x growth)."""
    sx = np.asarray(sx, dtype=np.float64)
    hop = 1.0 - overlap
    factor = np.asarray(csf, dtype=np.float64) * hop + overlap
    sw = np.ceil(sx / factor)
    swold = sw
    while np.any(np.floor(sw * overlap) != sw * overlap):
        sw = sw + 1
        if np.any(sw > 2 * swold):
            sw = swold
            break
    noverlap = np.floor(sw * overlap)
    nhop = sw - noverlap
    return sw.astype(np.int64), nhop.astype(np.int64)


class Win2WinAux:
    """imagetools.py:398-454 (``win2winaux``) in FP64.

    Attributes mirror the reference: sx, sw, csf, nhop, ws (P,sw0,sw1),
    nom, denom, w (the OLA of all windows, size sxp).
    """

    def __init__(self, sx, csf, overlap, windowfun='barthann'):
        sw, nhop = geometry(sx, csf, overlap)
        csf = tuple(int(c) for c in csf)
        P = csf[0] * csf[1]
        w = window(sw)  # imagetools.py:438 hard-codes 'barthann'
        nom = np.tile(w, (P, 1, 1))
        wimg = ola(nom, csf, sw, nhop)                    # imagetools.py:442
        denom = chop_pad(wimg, csf, sw, nhop, sw)          # imagetools.py:443
        self.sx = np.asarray(sx, dtype=np.int64)
        self.sw = sw
        self.csf = csf
        self.nhop = nhop
        self.overlap = overlap
        self.nom = nom
        self.denom = denom
        self.w = wimg
        self.ws = nom / denom                              # imagetools.py:445

    @property
    def sxp(self):
        return self.nhop * np.asarray(self.csf) + self.sw - self.nhop


def sxp_of(csf, sw, nhop):
    """gputools.py:375 ``sxp = nhop*csf + sw - nhop``."""
    return np.asarray(nhop) * np.asarray(csf) + np.asarray(sw) - np.asarray(nhop)


# --------------------------------------------------------------------------
# Index primitives (gputools.cu)
# --------------------------------------------------------------------------
def pad(x, sz, offset=(0, 0)):
    """gputools.cu:154-166 zeroPadKernel (via pad_cpu2gpu, gputools.py:1099-1137):
    out[i,j] = x[i-s, j-t] inside, 0 elsewhere; anything past ``sz`` is cut."""
    sz = (int(sz[0]), int(sz[1]))
    s, t = int(offset[0]), int(offset[1])
    out = np.zeros(sz, dtype=np.result_type(x, np.float64))
    r1 = min(sz[0], x.shape[0] + s)
    c1 = min(sz[1], x.shape[1] + t)
    if r1 > s and c1 > t:
        out[s:r1, t:c1] = x[: r1 - s, : c1 - t]
    return out


def patch_origins(csf, nhop):
    """Patch k = m*csf1 + n starts at (m*nhop0, n*nhop1) (gputools.cu:111-117)."""
    return [(m * int(nhop[0]), n * int(nhop[1]))
            for m in range(int(csf[0])) for n in range(int(csf[1]))]


def chop_pad(x, csf, sw, nhop, sz, offset=(0, 0)):
    """gputools.py:364-401 -> gputools.cu:105-126 (chop_pad_(Complex)Kernel).

    x is first zero-padded (or cut) to sxp, then patch k = xp[p_k : p_k+sw] is
    placed at ``offset`` inside a zero tile of size ``sz``.
    """
    sxp = sxp_of(csf, sw, nhop)
    xp = pad(x, sxp)
    sz = (int(sz[0]), int(sz[1]))
    o0, o1 = int(offset[0]), int(offset[1])
    P = int(csf[0]) * int(csf[1])
    out = np.zeros((P, sz[0], sz[1]), dtype=xp.dtype)
    h = min(int(sw[0]), sz[0] - o0)
    w = min(int(sw[1]), sz[1] - o1)
    for k, (p0, p1) in enumerate(patch_origins(csf, nhop)):
        out[k, o0:o0 + h, o1:o1 + w] = xp[p0:p0 + h, p1:p1 + w]
    return out


def chop_mod_pad(x, ws, csf, sw, nhop, sz, offset=(0, 0)):
    """gputools.py:322-361 -> gputools.cu:37-58: chop x multiplied by ws_k."""
    sw = (int(sw[0]), int(sw[1]))
    tiles = chop_pad(x, csf, sw, nhop, sw)
    tiles = tiles * ws
    return pad_stack(tiles, sz, offset)


def pad_stack(x, sz, offset=(0, 0)):
    """gputools.py:1140-1181 -> gputools.cu:184-222: place each slice at
    ``offset`` in a zero tile of size ``sz``."""
    sz = (int(sz[0]), int(sz[1]))
    o0, o1 = int(offset[0]), int(offset[1])
    out = np.zeros((x.shape[0], sz[0], sz[1]), dtype=np.result_type(x, np.float64))
    h = min(x.shape[1], sz[0] - o0)
    w = min(x.shape[2], sz[1] - o1)
    out[:, o0:o0 + h, o1:o1 + w] = x[:, :h, :w]
    return out


def crop_stack(x, sz, offset=(0, 0)):
    """gputools.py:599-640 -> gputools.cu:225-251."""
    o0, o1 = int(offset[0]), int(offset[1])
    return x[:, o0:o0 + int(sz[0]), o1:o1 + int(sz[1])].copy()


def crop(x, sz, offset=(0, 0)):
    """gputools.py:546-569 -> gputools.cu:254-274 (offset must keep the crop
    inside ``x``; the reference leaves such pixels uninitialised)."""
    o0, o1 = int(offset[0]), int(offset[1])
    s0, s1 = int(sz[0]), int(sz[1])
    if o0 < 0 or o1 < 0 or o0 + s0 > x.shape[0] or o1 + s1 > x.shape[1]:
        raise ValueError('crop window outside the array')
    return x[o0:o0 + s0, o1:o1 + s1].copy()


def ola(xs, csf, wm, nhop, offset=(0, 0)):
    """gputools.py:1068-1096 -> gputools.cu:276-314 (ola_Kernel_test).

    Output size ``nhop*csf + wm - nhop``; pixel (i,j) sums patch k's element at
    ``offset + (i,j) - p_k`` over every patch whose extent ``wm`` covers it.
    The reference reads past a slice when ``offset+wm`` exceeds the slice
    (the F.cnvtp call, olaGPU.py:307-309); this restatement refuses that case.
    """
    wm = (int(wm[0]), int(wm[1]))
    om, on = int(offset[0]), int(offset[1])
    if om + wm[0] > xs.shape[1] or on + wm[1] > xs.shape[2]:
        raise ValueError('OLA extent exceeds the patch stack slices')
    sx = sxp_of(csf, wm, nhop)
    out = np.zeros((int(sx[0]), int(sx[1])), dtype=np.result_type(xs, np.float64))
    for k, (p0, p1) in enumerate(patch_origins(csf, nhop)):
        out[p0:p0 + wm[0], p1:p1 + wm[1]] += xs[k, om:om + wm[0], on:on + wm[1]]
    return out


def ola_flat(xs, csf, wm, nhop):
    """gputools.cu:295-314 (ola_Kernel_test) as the reference runs it when the
    extent wm exceeds the slices (F.cnvtp passes sw+sf-1 over sw-sized
    slices, olaGPU.py:307-309): element (r, c) of patch k is read from the
    flat stack at k*sh*sw + r*sw + c, so columns past the slice wrap into
    the next row and rows past it into the next patch.  Reads past the end
    of the stack (undefined memory in the reference) read 0 here."""
    P, sh, sw = xs.shape
    flat = np.asarray(xs, dtype=np.float64).ravel()
    wm = (int(wm[0]), int(wm[1]))
    sx = sxp_of(csf, wm, nhop)
    out = np.zeros((int(sx[0]), int(sx[1])))
    r = np.arange(wm[0])[:, None]
    c = np.arange(wm[1])[None, :]
    for k, (p0, p1) in enumerate(patch_origins(csf, nhop)):
        idx = k * sh * sw + r * sw + c
        vals = np.where(idx < flat.size, flat[np.minimum(idx, flat.size - 1)], 0.0)
        out[p0:p0 + wm[0], p1:p1 + wm[1]] += vals
    return out


# --------------------------------------------------------------------------
# The operator (olaGPU.py)
# --------------------------------------------------------------------------
MODES = ('valid', 'same', 'full', 'circ')


def out_size(sx, sf, mode):
    """olaGPU.py:127-136."""
    sx = np.asarray(sx)
    sf = np.asarray(sf)
    if mode == 'valid':
        return sx - sf + 1
    if mode == 'same':
        return sx.copy()
    if mode == 'full':
        return sx + sf - 1
    if mode == 'circ':
        return sx.copy()
    raise NotImplementedError('Not a valid mode!')


def crop_offset(sf, mode):
    """Offset of the cropped output inside the OLA buffer (olaGPU.py:239-244)."""
    sf = np.asarray(sf)
    if mode == 'valid':
        return sf - 1
    if mode == 'same':
        return np.floor(sf / 2).astype(np.int64)
    if mode == 'full':
        return np.zeros(2, dtype=np.int64)
    raise NotImplementedError('Not a valid mode!')


def fft_tile(sw, sf):
    """olaGPU.py:139-141: T = 2**ceil(log2(sw+sf-1)) per dimension."""
    s = np.asarray(sw) + np.asarray(sf) - 1
    return tuple(int(v) for v in 2 ** np.ceil(np.log2(s)))


# FFT execution settings (the same FP64 DFT either way).  WORKERS = 1 runs
# numpy's pocketfft on one thread; WORKERS > 1 runs scipy.fft's pocketfft on
# that many threads (the large parity cases and the all-cores CPU baseline).
# SINGLE = True runs the transforms, and with them the spectral products, in
# complex64 (the FP32 CPU-baseline variant; overlap-add sums stay FP64).
WORKERS = 1
SINGLE = False


def set_fft(workers=1, single=False):
    """Select the FFT thread count and precision; returns the old setting."""
    global WORKERS, SINGLE
    old = (WORKERS, SINGLE)
    WORKERS, SINGLE = max(1, int(workers)), bool(single)
    return old


def _fft(a, inverse):
    if SINGLE:
        a = np.asarray(a, dtype=np.complex64)
    if WORKERS > 1 or SINGLE:
        import scipy.fft as sf
        fn = sf.ifft2 if inverse else sf.fft2
        return fn(a, axes=(-2, -1), workers=WORKERS)
    return (np.fft.ifft2 if inverse else np.fft.fft2)(a, axes=(-2, -1))


def fft2(a):
    return _fft(a, False)


def ifft2(a):
    return _fft(a, True)


class OlaOracle:
    """olaGPU.py:11-311 ``OlaGPU`` in FP64.

    ``OlaOracle(f, sx, mode, winaux)`` with a (P,sf0,sf1) PSF stack -> 'F'
    instance (caches F^); ``OlaOracle(x, sf, mode, winaux)`` with a 2-D image
    -> 'X' instance (caches the spectra of the windowed patches).
    ``tile`` overrides the FFT tile (any tile >= sw+sf-1 gives the same result).
    """

    def __init__(self, z, sx, mode, winaux, tile=None):
        z = np.asarray(z, dtype=np.float64)
        sf = np.array(z.shape[-2:])
        sx = np.array(sx)
        sw = np.asarray(winaux.sw)
        csf = winaux.csf
        if z.ndim == 3 and np.all(sx > sf):            # olaGPU.py:104-111
            self.f, self.x, self.id = z, None, 'F'
            if z.shape[0] != int(np.prod(csf)):
                raise IOError('Size missmatch between winaux and PSF size!')
        elif z.ndim == 2 and np.all(sx < sf):           # olaGPU.py:113-118
            self.f, self.x, self.id = None, z, 'X'
            sf = sx
            sx = np.array(z.shape)
        else:                                           # olaGPU.py:120-121
            raise IOError('Size missmatch')
        if np.any(np.asarray(winaux.sx) != sx):         # olaGPU.py:124-125
            raise IOError('Size missmatch between winaux and image size!')
        self.sy = out_size(sx, sf, mode)
        self.mode = mode
        self.sf = sf
        self.sx = sx
        self.winaux = winaux
        self.sfft = sw + sf - 1
        self.T = tile if tile is not None else fft_tile(sw, sf)
        if self.id == 'F':
            self.spec = fft2(pad_stack(self.f, self.T))
            self.sz = sx
        else:
            self.spec = fft2(chop_mod_pad(self.x, winaux.ws, csf, sw,
                                          winaux.nhop, self.T))
            self.sz = sf

    # -- forward ------------------------------------------------------------
    def cnv(self, u):
        """olaGPU.py:170-257: y = crop(OLA_{sw+sf-1}(Re IFFT(C^ . U^)))."""
        u = np.asarray(u, dtype=np.float64)
        wa = self.winaux
        if u.ndim == 3 and self.id == 'X':
            if u.shape[0] != int(np.prod(wa.csf)):
                raise IOError('Size missmatch between winaux and PSF size!')
            U = fft2(pad_stack(u, self.T))
        elif u.ndim == 2 and self.id == 'F':
            if np.any(np.array(u.shape) != np.asarray(wa.sx)):
                raise IOError('Size missmatch between winaux and image size!')
            U = fft2(chop_pad(u, wa.csf, wa.sw, wa.nhop, self.T)
                     * pad_stack(wa.ws, self.T))
        else:
            raise IOError('Size missmatch')
        ys = np.real(ifft2(self.spec * U))
        ext = np.asarray(wa.sw) + self.sf - 1
        yo = ola(ys, wa.csf, ext, wa.nhop)
        if self.mode == 'circ':
            raise NotImplementedError('Not a valid mode!')
        return crop(yo, self.sy, crop_offset(self.sf, self.mode))

    # -- transposes ---------------------------------------------------------
    def residual_tiles(self, y, exact):
        """Residual patches G_k placed in the FFT tile.

        exact=False: olaGPU.py:276-286 -- chop y (padded to sxp) with patch
        size sw at tile offset oc (= sf-1 valid / floor(sf/2) same / 0 full;
        the reference stops at valid/same, full extends the same rule).
        exact=True : the full correlation support, tile b in [0, sw+sf-1)
        holds y[p_k + b - oc] (zero outside the image).
        """
        wa = self.winaux
        if self.mode not in ('valid', 'same', 'full'):
            raise NotImplementedError('Not a valid mode!')
        oc = crop_offset(self.sf, self.mode)
        y = np.asarray(y, dtype=np.float64)
        if not exact:
            return chop_pad(y, wa.csf, wa.sw, wa.nhop, self.T, oc)
        ext = np.asarray(wa.sw) + self.sf - 1
        P = int(np.prod(wa.csf))
        out = np.zeros((P, self.T[0], self.T[1]))
        for k, (p0, p1) in enumerate(patch_origins(wa.csf, wa.nhop)):
            for b0 in range(int(ext[0])):
                i = p0 + b0 - int(oc[0])
                if i < 0 or i >= y.shape[0]:
                    continue
                j0 = p1 - int(oc[1])
                lo = max(0, -j0)
                hi = min(int(ext[1]), y.shape[1] - j0)
                if hi > lo:
                    out[k, b0, lo:hi] = y[i, j0 + lo:j0 + hi]
        return out

    def cnvtp(self, y, exact=False, k1=False):
        """olaGPU.py:261-311: Z = IFFT(G^ . conj(C^)).

        'X' -> Re Z[:, :sf, :sf] (PSF-shaped); 'F' -> crop to sw, multiply by
        the windows, OLA with extent sw (the reference passes sw+sf-1 here,
        reading past each slice; fixed), crop to sx.  k1=True reproduces the
        reference's read past the slices (ola_flat) instead.
        """
        wa = self.winaux
        G = fft2(self.residual_tiles(y, exact))
        Z = np.real(ifft2(G * np.conj(self.spec)))
        if self.id == 'X':
            return crop_stack(Z, self.sz)
        zs = crop_stack(Z, wa.sw) * wa.ws
        if k1:
            zc = ola_flat(zs, wa.csf, np.asarray(wa.sw) + self.sf - 1, wa.nhop)
        else:
            zc = ola(zs, wa.csf, wa.sw, wa.nhop)
        return crop(zc, self.sx)


# --------------------------------------------------------------------------
# Functional shortcuts used by the tests and the CPU baseline
# --------------------------------------------------------------------------
def forward(x, f, winaux, mode='valid', tile=None):
    """B(x, f) = F x = X f."""
    return OlaOracle(f, np.shape(x), mode, winaux, tile).cnv(x)


def adjoint_x(g, f, winaux, sx, mode='valid', exact=True, tile=None):
    """F' g (gradient of <B(x,f), g> with respect to x)."""
    return OlaOracle(f, sx, mode, winaux, tile).cnvtp(g, exact=exact)


def adjoint_f(g, x, winaux, sf, mode='valid', exact=True, tile=None):
    """X' g (gradient of <B(x,f), g> with respect to f)."""
    return OlaOracle(x, sf, mode, winaux, tile).cnvtp(g, exact=exact)
