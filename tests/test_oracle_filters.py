"""Pins of the FP64 filter and deconvolution oracles (CPU only): library
routines (np.diff, scipy.ndimage), brute-force transcriptions of the CUDA
index logic, grid-search minimisation and closed-form Fourier divisions."""
import numpy as np
import pytest
import scipy.ndimage as nd
import scipy.signal as ss

from oracle import deconv_oracle as DO
from oracle import filters_oracle as FO


def test_gradient_vs_diff():
    y = np.random.default_rng(0).random((7, 9))
    gx, gy = FO.gradient(y, 'valid')
    np.testing.assert_allclose(gx, -np.diff(y, axis=1) / np.sqrt(2), rtol=1e-15)
    np.testing.assert_allclose(gy, -np.diff(y, axis=0) / np.sqrt(2), rtol=1e-15)
    sx, sy = FO.gradient(y, 'same')
    assert sx.shape == y.shape and np.all(sx[:, -1] == 0) and np.all(sy[-1] == 0)


def test_laplace_vs_ndimage():
    y = np.random.default_rng(1).random((10, 12))
    k = np.array([[0, -1, 0], [-1, 4, -1], [0, -1, 0]]) / np.sqrt(20)
    ref = nd.correlate(y, k)[1:-1, 1:-1]
    np.testing.assert_allclose(FO.laplace(y, 'valid'), ref, rtol=1e-13)
    s = FO.laplace(y, 'same')
    np.testing.assert_allclose(s[1:-1, 1:-1], ref, rtol=1e-13)
    assert np.all(s[0] == 0) and np.all(s[:, -1] == 0)
    st = FO.laplace_stack(np.stack([y, y, y]))
    assert np.all(st[0] == 0) and np.all(st[2] == 0)
    np.testing.assert_allclose(st[1], s)


def test_laplace3d_vs_ndimage():
    y = np.random.default_rng(2).random((5, 6, 7))
    k = np.zeros((3, 3, 3))
    k[1, 1, 1] = 6
    for a, b, c in [(0, 1, 1), (2, 1, 1), (1, 0, 1), (1, 2, 1), (1, 1, 0), (1, 1, 2)]:
        k[a, b, c] = -1
    ref = nd.correlate(y, k)
    np.testing.assert_allclose(FO.laplace3d(y)[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1], rtol=1e-13)


def test_impad_vs_kernel_loop():
    y = np.random.default_rng(3).random((5, 6))
    for sf in [(4, 5), (19, 19), (3, 2)]:
        out = FO.impad(y, sf)
        top, left = sf[0] // 2, sf[1] // 2
        for i in range(out.shape[0]):
            for j in range(out.shape[1]):
                si = 0 if i < top else (y.shape[0] - 1 if i - top >= y.shape[0] else i - top)
                sj = 0 if j < left else (y.shape[1] - 1 if j - left >= y.shape[1] else j - left)
                assert out[i, j] == y[si, sj]


def _shock_cuda_pixel(a, i, j, dt, h):
    """kernel.cu:15-86 for one pixel, reading the halo exactly as the
    shared-memory loader does."""
    H, W = a.shape

    def s(di, dj):
        ii, jj = i + di, j + dj
        ii = 0 if ii < 0 else (H if ii >= H else ii)
        jj = 0 if jj < 0 else (W if jj >= W else jj)
        return a[ii, jj] if (ii < H and jj < W) else 0.0

    t1 = -s(0, 0) + s(0, 1)
    t2 = s(0, 0) - s(0, -1)
    Ix = (t1 + t2) / 2
    t3 = 0.0 if t1 * t2 < 0 else min(abs(t1), abs(t2))
    t1 = -s(0, 0) + s(1, 0)
    t2 = s(0, 0) - s(-1, 0)
    Iy = (t1 + t2) / 2
    flow = 0.0 if t1 * t2 < 0 else min(abs(t1), abs(t2))
    flow = np.sqrt(t3 * t3 + flow * flow)
    t1 = s(0, 1) + s(0, -1) - 2 * s(0, 0)
    t2 = s(1, 0) + s(-1, 0) - 2 * s(0, 0)
    t3 = (s(1, 1) - s(1, -1) - s(-1, 1) + s(-1, -1)) / 4
    Inn = (t2 * Iy * Iy + 2 * t3 * Iy * Ix + t1 * Ix * Ix) / (Ix * Ix + Iy * Iy + 1e-8)
    if Ix == 0 and Iy == 0:
        Inn = t1
    return s(0, 0) - dt * np.copysign(flow, Inn) / h


def test_shock_step_vs_cuda_transcription():
    a = np.random.default_rng(4).random((9, 11))
    a[3:6, 2:8] += 1.0
    out = FO.shock_step(a, 0.1, 1.0)
    ref = np.array([[_shock_cuda_pixel(a, i, j, 0.1, 1.0) for j in range(11)] for i in range(9)])
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-15)


def test_bilateral_bruteforce():
    a = np.random.default_rng(5).random((8, 10))
    w, sd, sr = 2, 1.5, 0.2
    out = FO.bilateral(a, w, sd, sr)
    ref = np.zeros_like(a)
    for i in range(8):
        for j in range(10):
            num = den = 0.0
            for x in range(-w, w + 1):
                for y in range(-w, w + 1):
                    if 0 <= i + y < 8 and 0 <= j + x < 10:
                        v = a[i + y, j + x]
                        wt = np.exp(-((v - a[i, j]) ** 2 / (2 * sr * sr) + (x * x + y * y) / (2 * sd * sd)))
                        num += wt * v
                        den += wt
            ref[i, j] = num / den
    np.testing.assert_allclose(out, ref, rtol=1e-13)


def test_shrink23_grid_minimum():
    """Away from the zero/nonzero switch the 4 Newton steps land on the global
    minimiser of |x|^(2/3) + beta/2 (x - v)^2 (grid search)."""
    rng = np.random.default_rng(6)
    for beta in (1.0, 8.0, 64.0):
        v = rng.uniform(-2, 2, 40)
        x = FO.shrink23(v, beta)
        grid = np.linspace(-3, 3, 600001)
        for vi, xi in zip(v, x):
            obj = np.abs(grid) ** (2 / 3) + beta / 2 * (grid - vi) ** 2
            xb = grid[np.argmin(obj)]
            ob = obj.min()
            o0 = beta / 2 * vi * vi
            if abs(ob - o0) > 1e-3 * max(1.0, o0):   # not at the switch point
                assert abs(xi - xb) < 2e-3 or (xi == 0 and abs(xb) < 2e-3), (beta, vi, xi, xb)


def test_resize_vs_map_coordinates():
    a = np.random.default_rng(7).random((13, 17))
    for osz in [(26, 34), (7, 9), (13, 17), (20, 11)]:
        ii, jj = np.meshgrid(np.arange(osz[0]), np.arange(osz[1]), indexing='ij')
        coords = [ii / osz[0] * 13 - 0.5, jj / osz[1] * 17 - 0.5]
        ref = nd.map_coordinates(a, coords, order=1, mode='nearest')
        np.testing.assert_allclose(FO.resize(a, osz), ref, rtol=1e-12, atol=1e-14)


def test_edgetaper_tables_and_loop():
    wx, wy = FO.edgetaper_tables((6, 9))          # made odd: 7, 9
    gx = ss.get_window('barthann', 9)
    gy = ss.get_window('barthann', 7)
    np.testing.assert_allclose(wx, gx[:4] / (gx.max() * gy.max()))
    np.testing.assert_allclose(wy, gy[:3] / (gx.max() * gy.max()))
    y = np.ones((10, 12))
    out = FO.edgetaper(y, (6, 9))
    assert out[5, 6] == 1.0                        # interior untouched
    np.testing.assert_allclose(out[0, 6], wy[0])
    np.testing.assert_allclose(out[9, 11], wy[0] * wx[0])
    np.testing.assert_allclose(out[8, 1], wy[1] * wx[1])


def test_fourier_divisions_closed_forms():
    rng = np.random.default_rng(8)
    F = rng.standard_normal((3, 8, 8)) + 1j * rng.standard_normal((3, 8, 8))
    U = rng.standard_normal((3, 8, 8)) + 1j * rng.standard_normal((3, 8, 8))
    L = np.abs(np.fft.fft2(np.pad([[-1.0, 1.0]], ((0, 7), (0, 6))))) ** 2
    np.testing.assert_allclose(DO.fourier_div(F, F * U, L, 0.0, 0.0), U, rtol=1e-12)
    # Tikhonov: z minimises |F z - Y|^2 + beta |z|^2 + alpha L |z|^2 per frequency
    Y = F * U
    z = DO.fourier_div(F, Y, L, 0.3, 0.1)
    np.testing.assert_allclose((np.abs(F) ** 2 + 0.3 * L + 0.1) * z, np.conj(F) * Y, rtol=1e-12)
    Xx, Xy = F, np.roll(F, 1)
    Z = DO.fourier_div_g(Xx, Xy, Xx * U, Xy * U, L, 0.0, 0.0)
    np.testing.assert_allclose(Z, U, rtol=1e-12)
    gx = np.fft.fft2(np.pad([[-1.0, 1.0]], ((0, 7), (0, 6))))
    gy = np.fft.fft2(np.pad([[-1.0], [1.0]], ((0, 6), (0, 7))))
    Ls = np.abs(gx) ** 2 + np.abs(gy) ** 2
    a, b, g = 0.5, 4.0, 1e-4
    wx, wy = rng.standard_normal(F.shape) + 0j, rng.standard_normal(F.shape) + 0j
    ftpy = np.conj(F) * Y
    Zs = DO.fourier_div_s(gx, gy, wx, wy, ftpy, F, Ls, a, b, g)
    np.testing.assert_allclose((np.abs(F) ** 2 / (a * b) + Ls + g) * Zs,
                               np.conj(gx) * wx + np.conj(gy) * wy + ftpy / (a * b), rtol=1e-11)
