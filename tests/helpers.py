"""Shared test helpers: seeded inputs and error metrics."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def rand_psfs(rng, P, sf, nonneg=True):
    f = rng.random((P, sf[0], sf[1])) if nonneg else rng.standard_normal((P, sf[0], sf[1]))
    if nonneg:
        f /= f.sum(axis=(1, 2), keepdims=True)
    return f


def case_inputs(seed, sx, csf, sf, mode='valid'):
    """x ~ U[0,1], nonnegative PSFs summing to one, residual g ~ N(0,1) of the
    forward's output size."""
    rng = np.random.default_rng(seed)
    x = rng.random(sx)
    P = csf[0] * csf[1]
    f = rand_psfs(rng, P, sf)
    sx = np.asarray(sx)
    sfa = np.asarray(sf)
    sy = {'valid': sx - sfa + 1, 'same': sx, 'full': sx + sfa - 1}[mode]
    g = rng.standard_normal(tuple(int(v) for v in sy))
    return x, f, g


# (sx, csf, sf, overlap) geometries exercised by the parity tests
GEOMS = [
    ((64, 64), (2, 2), (5, 5), 0.5),
    ((96, 96), (3, 3), (9, 9), 0.5),
    ((70, 90), (2, 3), (7, 5), 0.5),
    ((256, 256), (4, 4), (15, 15), 0.5),
    ((120, 100), (3, 4), (6, 8), 0.5),
    ((200, 180), (4, 3), (11, 13), 0.25),
    ((48, 40), (1, 1), (3, 3), 0.5),
    # PSF rows beyond three slots per thread (17 > 3 * 64/16): the column
    # passes' unpruned transforms
    ((60, 60), (3, 3), (17, 17), 0.5),
]
