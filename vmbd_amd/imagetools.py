"""Host-side image utilities of the drop-in (imagetools.py of the reference).

Window and geometry math runs in float64 on the host; the float32 window
stack is moved to the GPU once.  The kernels read the windows from a small
table of distinct windows (interior patches share one window; only border
patches differ) indexed per patch.
"""
import hashlib

import numpy as np

WINDOW_FLOOR = 1e-16  # imagetools.py:342


def get_window_barthann(n):
    """Periodic Bartlett-Hann window of length n, i.e. the symmetric window
    of length n+1 without its last sample (what scipy.signal.get_window
    returns for 'barthann'; used by imagetools.py:332 and gputools.py:653)."""
    n = int(n)
    if n <= 1:
        return np.ones(n)
    m = n + 1
    k = np.arange(m, dtype=np.float64)
    fac = np.abs(k / (m - 1.0) - 0.5)
    w = 0.62 - 0.48 * fac + 0.38 * np.cos(2.0 * np.pi * fac)
    return w[:n]


def window1d(n):
    """Symmetrised, min-shifted 1-D factor of the 2-D window
    (imagetools.py:332-337)."""
    w = get_window_barthann(n)
    w = (w + w[::-1]) / 2.0
    return w - w.min()


def window(sw, window='barthann'):
    """2-D patch window (imagetools.py:314-344): outer product of the 1-D
    factors, scaled to max 1, floored at 1e-16."""
    if window != 'barthann':
        raise NotImplementedError('only the barthann window is implemented')
    w1 = window1d(int(sw[0]))
    w2 = window1d(int(sw[1]))
    www = np.outer(w1, w2)
    www /= www.max()
    return np.maximum(www, WINDOW_FLOOR)


def window_geometry(sx, csf, overlap):
    """Window size and hop (imagetools.py:423-436).  Both dimensions grow
    together until sw*overlap is integral; after 2x growth the first guess is
    kept (then the hop differs from (1-overlap)*sw in that dimension)."""
    sx = np.asarray(sx, dtype=np.float64)
    hop = 1.0 - overlap
    factor = np.asarray(csf, dtype=np.float64) * hop + overlap
    sw = np.ceil(sx / factor)
    sw0 = sw
    while np.any(np.floor(sw * overlap) != sw * overlap):
        sw = sw + 1
        if np.any(sw > 2 * sw0):
            sw = sw0
            break
    nhop = sw - np.floor(sw * overlap)
    return sw.astype(np.int64), nhop.astype(np.int64)


def patch_origin(k, csf, nhop):
    m, n = divmod(int(k), int(csf[1]))
    return m * int(nhop[0]), n * int(nhop[1])


def normalized_windows(sx, csf, overlap):
    """ws_k = w / (sum of all windows covering the pixel), float64
    (imagetools.py:438-445).  Returns (sw, nhop, ws, nom2d, wimg)."""
    sw, nhop = window_geometry(sx, csf, overlap)
    w = window(sw)
    P = int(csf[0]) * int(csf[1])
    sxp = nhop * np.asarray(csf) + sw - nhop
    wimg = np.zeros((int(sxp[0]), int(sxp[1])))
    for k in range(P):
        p0, p1 = patch_origin(k, csf, nhop)
        wimg[p0:p0 + sw[0], p1:p1 + sw[1]] += w
    ws = np.empty((P, int(sw[0]), int(sw[1])))
    for k in range(P):
        p0, p1 = patch_origin(k, csf, nhop)
        ws[k] = w / wimg[p0:p0 + sw[0], p1:p1 + sw[1]]
    return sw, nhop, ws, w, wimg


def window_classes(ws32):
    """Distinct windows of a float32 stack: (table (n, sw0, sw1), idx (P,))."""
    keys = {}
    idx = np.empty(ws32.shape[0], dtype=np.int32)
    table = []
    for k in range(ws32.shape[0]):
        h = hashlib.blake2b(np.ascontiguousarray(ws32[k]).tobytes(), digest_size=16).digest()
        if h not in keys:
            keys[h] = len(table)
            table.append(ws32[k])
        idx[k] = keys[h]
    return np.ascontiguousarray(np.stack(table)), idx


class win2winaux:
    """Drop-in of imagetools.win2winaux (imagetools.py:398-454).

    win2winaux(sx, csf, overlap, windowfun='barthann').  Attributes: sx, sw,
    csf, nhop, ws_gpu, nom_gpu, denom_gpu, w_gpu (device tensors, created on
    first access), plus ``win_table``/``win_idx`` for the kernels.
    ``windowfun`` is accepted and ignored, as in the reference (:438).
    """

    def __init__(self, sx, csf, overlap, windowfun='barthann', device=None):
        sw, nhop, ws, w, wimg = normalized_windows(sx, csf, overlap)
        self.sx = np.asarray(sx).astype(np.int64)
        self.sw = sw
        self.csf = tuple(int(c) for c in csf)
        self.nhop = nhop
        self.overlap = overlap
        self.ws = ws.astype(np.float32)
        self._nom = w
        self._wimg = wimg
        self.win_table, self.win_idx = window_classes(self.ws)
        self._device = device
        self._cache = {}

    def window_factors(self):
        """(a, b, floor): the window is max(a[r] * b[c], floor) (the outer
        product of the 1-D factors scaled to max 1, as in ``window``)."""
        w1 = window1d(int(self.sw[0]))
        w2 = window1d(int(self.sw[1]))
        a = np.ascontiguousarray(w1 / w1.max(), dtype=np.float32)
        b = np.ascontiguousarray(w2 / w2.max(), dtype=np.float32)
        return a, b, WINDOW_FLOOR

    @property
    def P(self):
        return self.csf[0] * self.csf[1]

    @property
    def sxp(self):
        return self.nhop * np.asarray(self.csf) + self.sw - self.nhop

    def _dev(self, name, arr):
        import torch
        if name not in self._cache:
            dev = self._device if self._device is not None else 'cuda'
            self._cache[name] = torch.as_tensor(np.ascontiguousarray(arr, dtype=np.float32), device=dev)
        return self._cache[name]

    @property
    def ws_gpu(self):
        return self._dev('ws', self.ws)

    @property
    def nom_gpu(self):
        return self._dev('nom', np.tile(self._nom, (self.P, 1, 1)))

    @property
    def w_gpu(self):
        return self._dev('w', self._wimg)

    @property
    def denom_gpu(self):
        d = np.stack([self._wimg[p0:p0 + self.sw[0], p1:p1 + self.sw[1]]
                      for p0, p1 in (patch_origin(k, self.csf, self.nhop) for k in range(self.P))])
        return self._dev('denom', d)


def fft_tile(sw, sf, minimum=16):
    """olaGPU.py:139-141: T = 2**ceil(log2(sw+sf-1)), at least ``minimum``
    (any tile >= sw+sf-1 gives the same linear convolution)."""
    s = np.asarray(sw) + np.asarray(sf) - 1
    t = (2 ** np.ceil(np.log2(np.maximum(s, 1)))).astype(np.int64)
    return tuple(int(max(minimum, v)) for v in t)


# ------------------------------------------------------------ frame I/O
def imread(filename):
    """8-bit PNG frame as float32 in [0, 1] (what matplotlib's imread gives
    for the reference's frames, vmbd_chimney.py:55); colour images are
    reduced to their green channel (rgb2gray, imagetools.py:266-287)."""
    from PIL import Image
    a = np.asarray(Image.open(filename))
    if a.ndim > 2:
        a = a[..., 1]
    if a.dtype == np.uint8:
        return a.astype(np.float32) / np.float32(255.0)
    if a.dtype == np.uint16:
        return a.astype(np.float32) / np.float32(65535.0)
    return a.astype(np.float32)


def imwrite(x, filename):
    """imagetools.py:290-311: an image in [0, 1] written as 8 bits (x * 255
    truncated to an integer; values outside [0, 1] are clipped, where the
    reference's uint8 conversion would wrap them)."""
    from PIL import Image
    t = np.array(x, dtype=np.float64, copy=True)
    if t.dtype != np.uint8:
        t *= 255.0
    Image.fromarray(np.uint8(np.clip(t, 0, 255))).save(filename)


def gridF(fs, csf):
    """imagetools.py:211-238: the PSF stack (P, sf0, sf1) laid out as one
    image of csf0 x csf1 tiles, patch k = m*csf1 + n at tile (m, n) (the
    reference's placement for the square grids it supports)."""
    fs = np.asarray(fs)
    c0, c1 = int(csf[0]), int(csf[1])
    h, w = fs.shape[1], fs.shape[2]
    f = np.ones((h * c0, w * c1))
    for i in range(c0 * c1):
        k, l = divmod(i, c1)
        f[k * h:(k + 1) * h, l * w:(l + 1) * w] = fs[i]
    return f


def randomwalk(sf, N=None, rng=None):
    """imagetools.py:99-132: a random walk of N steps (default prod(sf)/100)
    from the centre, each step moving every coordinate by -1, 0 or +1,
    clipped to [1, min(sf)-1]; normalised to sum 1.  `rng` is a numpy
    Generator (the reference uses pylab's global generator)."""
    rng = np.random.default_rng() if rng is None else rng
    sf = tuple(int(v) for v in np.atleast_1d(sf))
    if N is None:
        N = int(np.floor(np.prod(sf) / 100))
    w = np.zeros(sf)
    loc = np.ceil(np.array(sf) / 2.0)
    w[tuple(loc.astype(int))] = 1.0
    for _ in range(int(N)):
        loc = loc + ((rng.random(len(sf)) - 0.5) * 2).round()
        loc = np.clip(loc, 1, min(sf) - 1)
        w[tuple(loc.astype(int))] += 1
    return w / w.sum()
