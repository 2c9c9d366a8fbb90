"""FP64 NumPy restatement of the reference's image filters (test
infrastructure: only tests/, smoke() and bench.py's cpu_baseline use it).

Cited per function: kernel.cu, edgetaper.cu, resize.cu, gputools.py.
Stencils are evaluated out of place (Jacobi), the semantics the HIP kernels
implement (the reference's shock/bilateral kernels race, kernel.cu:86, 134).
"""
import numpy as np

from .ola_oracle import barthann_periodic

SQRT2H = np.sqrt(0.5)


def shock_step(a, dt=0.1, h=1.0):
    """kernel.cu:9-88, one step.  Neighbours above/left of the image replicate
    the edge, neighbours below/right read 0 (kernel.cu:22-30)."""
    c, flow, Inn = shock_terms(a)
    return c - dt * np.copysign(flow, Inn) / h


def shock_terms(a):
    """The pixel values c, the minmod gradient magnitude |grad I| and the
    second directional derivative Inn of one shock step (kernel.cu:9-88);
    the step is c - dt sign(Inn) |grad I| / h."""
    a = np.asarray(a, dtype=np.float64)
    H, W = a.shape
    p = np.zeros((H + 2, W + 2))
    p[1:H + 1, 1:W + 1] = a
    p[0, 1:W + 1] = a[0]          # row -1 -> row 0
    p[1:H + 1, 0] = a[:, 0]       # col -1 -> col 0
    p[0, 0] = a[0, 0]
    p[0, W + 1] = 0.0             # (-1, W): column past the end reads 0
    p[H + 1, 0] = 0.0
    c = p[1:H + 1, 1:W + 1]
    l, r = p[1:H + 1, 0:W], p[1:H + 1, 2:W + 2]
    u, d = p[0:H, 1:W + 1], p[2:H + 2, 1:W + 1]
    t1, t2 = -c + r, c - l
    Ix = (t1 + t2) / 2
    mx = np.where(t1 * t2 < 0, 0.0, np.minimum(np.abs(t1), np.abs(t2)))
    t1, t2 = -c + d, c - u
    Iy = (t1 + t2) / 2
    my = np.where(t1 * t2 < 0, 0.0, np.minimum(np.abs(t1), np.abs(t2)))
    flow = np.sqrt(mx * mx + my * my)
    Ixx = r + l - 2 * c
    Iyy = d + u - 2 * c
    Ixy = (p[2:H + 2, 2:W + 2] - p[2:H + 2, 0:W] - p[0:H, 2:W + 2] + p[0:H, 0:W]) / 4
    Inn = (Iyy * Iy * Iy + 2 * Ixy * Iy * Ix + Ixx * Ix * Ix) / (Ix * Ix + Iy * Iy + 1e-8)
    Inn = np.where((Ix == 0) & (Iy == 0), Ixx, Inn)
    return c, flow, Inn


def shock_filter(a, iters=3, dt=0.1, h=1.0):
    """gputools.py:1282-1314."""
    for _ in range(iters):
        a = shock_step(a, dt, h)
    return a


def bilateral(a, w=5, sigma_d=3.0, sigma_r=0.1):
    """kernel.cu:90-136: weights exp(-dv^2/(2 sr^2) - (x^2+y^2)/(2 sd^2)) over
    the in-image neighbours."""
    a = np.asarray(a, dtype=np.float64)
    H, W = a.shape
    num = np.zeros_like(a)
    den = np.zeros_like(a)
    for dy in range(-w, w + 1):
        for dx in range(-w, w + 1):
            sh = np.full((H, W), np.nan)
            i0, i1 = max(0, -dy), min(H, H - dy)
            j0, j1 = max(0, -dx), min(W, W - dx)
            sh[i0:i1, j0:j1] = a[i0 + dy:i1 + dy, j0 + dx:j1 + dx]
            ok = ~np.isnan(sh)
            v = np.where(ok, sh, 0.0)
            wt = np.exp(-((v - a) ** 2 / (2 * sigma_r ** 2) + (dx * dx + dy * dy) / (2 * sigma_d ** 2)))
            wt = np.where(ok, wt, 0.0)
            num += wt * v
            den += wt
    return num / den


def gradient(y, mode='valid'):
    """kernel.cu:180-242."""
    y = np.asarray(y, dtype=np.float64)
    gx = SQRT2H * (y[:, :-1] - y[:, 1:])
    gy = SQRT2H * (y[:-1, :] - y[1:, :])
    if mode == 'same':
        gx = np.concatenate([gx, np.zeros((y.shape[0], 1))], axis=1)
        gy = np.concatenate([gy, np.zeros((1, y.shape[1]))], axis=0)
    return gx, gy


def laplace(y, mode='valid'):
    """kernel.cu:268-333: (4c - sum of 4 neighbours)/sqrt(20)."""
    y = np.asarray(y, dtype=np.float64)
    c = y[1:-1, 1:-1]
    v = (4 * c - (y[2:, 1:-1] + y[:-2, 1:-1] + y[1:-1, 2:] + y[1:-1, :-2])) / np.sqrt(20.0)
    if mode == 'valid':
        return v
    out = np.zeros_like(y)
    out[1:-1, 1:-1] = v
    return out


def laplace_stack(y):
    """kernel.cu:335-375 as computed: per-slice 'same' Laplacian, first and
    last slice zero."""
    out = np.stack([laplace(s, 'same') for s in y])
    out[0] = 0
    out[-1] = 0
    return out


def laplace3d(y):
    """kernel.cu:377-418."""
    y = np.asarray(y, dtype=np.float64)
    out = np.zeros_like(y)
    c = y[1:-1, 1:-1, 1:-1]
    out[1:-1, 1:-1, 1:-1] = 6 * c - (y[2:, 1:-1, 1:-1] + y[:-2, 1:-1, 1:-1] + y[1:-1, 2:, 1:-1] +
                                     y[1:-1, :-2, 1:-1] + y[1:-1, 1:-1, 2:] + y[1:-1, 1:-1, :-2])
    return out


def impad(y, sf):
    """kernel.cu:244-266: replicate padding, floor(sf/2) before."""
    sf = np.broadcast_to(np.asarray(sf, dtype=int), (2,))
    top, left = sf[0] // 2, sf[1] // 2
    return np.pad(np.asarray(y, dtype=np.float64), ((top, sf[0] - top), (left, sf[1] - left)), mode='edge')


def shrink23(v, beta):
    """kernel.cu:555-578 in FP64: 4 Newton steps on |x|^(2/3) + beta/2 (x-v)^2,
    then the x = 0 test."""
    v = np.asarray(v, dtype=np.float64)
    x = v.copy()
    with np.errstate(divide='ignore', invalid='ignore', over='ignore'):
        for _ in range(4):
            t1 = (2.0 / 3.0) / np.cbrt(np.abs(x))
            x = x - (np.copysign(t1, x) + beta * (x - v)) / (t1 / (-3.0 * np.abs(x)) + beta)
        x = np.where(np.isfinite(x), x, 0.0)
        t = np.cbrt(np.abs(x)) ** 2
        x = np.where(t + beta * 0.5 * (x - v) ** 2 > beta * 0.5 * v * v, 0.0, x)
    return x


def shrink(v, beta, alpha):
    """kernel.cu:529-553 in FP64."""
    v = np.asarray(v, dtype=np.float64)
    x = v.copy()
    with np.errstate(divide='ignore', invalid='ignore', over='ignore'):
        for _ in range(4):
            fd = alpha * np.where(x > 0, 1.0, -1.0) * np.abs(x) ** (alpha - 1) + beta * (x - v)
            fdd = alpha * (alpha - 1) * np.abs(x) ** (alpha - 2) + beta
            x = x - fd / fdd
        x = np.where(np.isfinite(x), x, 0.0)
        x = np.where(np.abs(x) ** alpha + beta * 0.5 * (x - v) ** 2 > beta * 0.5 * v * v, 0.0, x)
    return x


def edgetaper_tables(sf):
    """gputools.py:651-659."""
    sf = np.broadcast_to(np.asarray(sf, dtype=int), (2,)).copy()
    sf = sf + (1 - np.mod(sf, 2))
    wx = barthann_periodic(sf[1])
    wy = barthann_periodic(sf[0])
    maxw = wx.max() * wy.max()
    hsf = sf // 2
    return wx[:hsf[1]] / maxw, wy[:hsf[0]] / maxw


def edgetaper(y, sf):
    """edgetaper.cu:6-31."""
    y = np.array(y, dtype=np.float64)
    wx, wy = edgetaper_tables(sf)
    H, W = y.shape
    hy, hx = len(wy), len(wx)
    # top rows take precedence over bottom rows when they overlap (if/else if)
    fy2 = np.ones(H)
    for i in range(H):
        if i < hy:
            fy2[i] = wy[i]
        elif i > H - 1 - hy:
            fy2[i] = wy[H - 1 - i]
    fx2 = np.ones(W)
    for j in range(W):
        if j < hx:
            fx2[j] = wx[j]
        elif j > W - 1 - hx:
            fx2[j] = wx[W - 1 - j]
    return y * fy2[:, None] * fx2[None, :]


def resize(a, out_shape):
    """resize.cu:5-15 with exact bilinear weights: sample at normalised
    (j/OW, i/OH), texel centres at +0.5, clamped addressing."""
    a = np.asarray(a, dtype=np.float64)
    H, W = a.shape
    OH, OW = int(out_shape[0]), int(out_shape[1])
    ys = np.arange(OH) / OH * H - 0.5
    xs = np.arange(OW) / OW * W - 0.5
    y0 = np.floor(ys)
    x0 = np.floor(xs)
    ay = ys - y0
    ax = xs - x0
    yi0 = np.clip(y0.astype(int), 0, H - 1)
    yi1 = np.clip(y0.astype(int) + 1, 0, H - 1)
    xi0 = np.clip(x0.astype(int), 0, W - 1)
    xi1 = np.clip(x0.astype(int) + 1, 0, W - 1)
    t0 = a[yi0][:, xi0] * (1 - ax) + a[yi0][:, xi1] * ax
    t1 = a[yi1][:, xi0] * (1 - ax) + a[yi1][:, xi1] * ax
    return t0 * (1 - ay[:, None]) + t1 * ay[:, None]
