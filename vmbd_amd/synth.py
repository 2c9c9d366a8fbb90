"""Seeded synthetic inputs for the benchmark and the tests (no datasets are
available offline).

* ``piecewise_smooth``: random axis-aligned rectangles of constant level plus
  Gaussian blobs, clipped to [0, 1].
* ``motion_psfs``: motion-streak PSFs from a random walk started at the centre
  (the scheme of imagetools.py:99-132), normalised to sum 1.
* ``gaussian_psfs``: anisotropic Gaussians (sigma in [0.8, 3], random angle),
  nonnegative, sum 1.
"""
import numpy as np


def piecewise_smooth(shape, rng, n_rect=20, n_blob=10):
    H, W = int(shape[0]), int(shape[1])
    x = np.full((H, W), 0.2 + 0.1 * rng.random())
    for _ in range(n_rect):
        h = int(rng.integers(max(2, H // 16), max(3, H // 3)))
        w = int(rng.integers(max(2, W // 16), max(3, W // 3)))
        i = int(rng.integers(0, max(1, H - h)))
        j = int(rng.integers(0, max(1, W - w)))
        x[i:i + h, j:j + w] = rng.random()
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    for _ in range(n_blob):
        ci, cj = rng.random() * H, rng.random() * W
        s = (0.02 + 0.08 * rng.random()) * min(H, W)
        a = rng.random() - 0.3
        x += a * np.exp(-((yy - ci) ** 2 + (xx - cj) ** 2) / (2 * s * s))
    return np.clip(x, 0.0, 1.0)


def motion_psf(sf, rng, steps=15):
    sf = (int(sf[0]), int(sf[1]))
    w = np.zeros(sf)
    loc = np.array([int(np.ceil(sf[0] / 2)) - 1, int(np.ceil(sf[1] / 2)) - 1])
    w[tuple(loc)] = 1.0
    hi = min(sf) - 1
    for _ in range(int(steps)):
        loc = loc + np.round((rng.random(2) - 0.5) * 2).astype(int)
        loc = np.clip(loc, 1, hi - 1) if hi > 1 else np.zeros(2, dtype=int)
        w[tuple(loc)] += 1.0
    return w / w.sum()


def motion_psfs(P, sf, rng, steps=15):
    return np.stack([motion_psf(sf, rng, steps) for _ in range(int(P))])


def gaussian_psf(sf, rng, smin=0.8, smax=3.0):
    sf = (int(sf[0]), int(sf[1]))
    sx_, sy_ = rng.uniform(smin, smax, 2)
    th = rng.uniform(0, np.pi)
    c0, c1 = (sf[0] - 1) / 2.0, (sf[1] - 1) / 2.0
    yy, xx = np.mgrid[0:sf[0], 0:sf[1]].astype(np.float64)
    u = (xx - c1) * np.cos(th) + (yy - c0) * np.sin(th)
    v = -(xx - c1) * np.sin(th) + (yy - c0) * np.cos(th)
    w = np.exp(-0.5 * ((u / sx_) ** 2 + (v / sy_) ** 2))
    return w / w.sum()


def gaussian_psfs(P, sf, rng):
    return np.stack([gaussian_psf(sf, rng) for _ in range(int(P))])
