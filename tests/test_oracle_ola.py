"""Pins of the FP64 oracle of the EFF/OLA operator (CPU only).

Every check below is independent of the oracle's own formula: scipy's window
and convolution routines, brute-force loops over the CUDA index expressions
of gputools.cu, closed-form invariants (partition of unity, delta PSF,
invariant PSF == conv2), adjoint dot-product identities and FP64 central
finite differences.
"""
import numpy as np
import pytest
import scipy.signal as ss

from oracle import ola_oracle as O
from tests.helpers import GEOMS, case_inputs, rand_psfs, rel_l2


# ---------------------------------------------------------------- windows
def test_barthann_matches_scipy_get_window():
    for n in range(1, 300):
        np.testing.assert_allclose(O.barthann_periodic(n), ss.get_window('barthann', n), rtol=0, atol=1e-15)


def test_window_construction_with_scipy():
    for sw in [(8, 8), (104, 104), (171, 228), (48, 40), (5, 7)]:
        w1 = ss.get_window('barthann', sw[0])
        w1 = (w1 + w1[::-1]) / 2
        w1 -= w1.min()
        w2 = ss.get_window('barthann', sw[1])
        w2 = (w2 + w2[::-1]) / 2
        w2 -= w2.min()
        ref = np.maximum(np.outer(w1, w2) / np.outer(w1, w2).max(), 1e-16)
        np.testing.assert_allclose(O.window(sw), ref, rtol=1e-14, atol=0)


# geometry values of the benchmark configurations (imagetools.py:423-436)
GEOM_TABLE = [
    ((256, 256), (4, 4), (104, 104), (52, 52), (260, 260)),
    ((512, 512), (8, 8), (114, 114), (57, 57), (513, 513)),
    ((1024, 1024), (10, 10), (188, 188), (94, 94), (1034, 1034)),
    ((2048, 2048), (16, 16), (242, 242), (121, 121), (2057, 2057)),
    ((768, 1024), (8, 8), (171, 228), (86, 114), (773, 1026)),   # bailed: odd sw0
    ((768, 1024), (6, 6), (220, 293), (110, 147), (770, 1028)),  # bailed: odd sw1
]


@pytest.mark.parametrize('sx,csf,sw,nhop,sxp', GEOM_TABLE)
def test_geometry_table(sx, csf, sw, nhop, sxp):
    s, h = O.geometry(sx, csf, 0.5)
    assert tuple(s) == sw and tuple(h) == nhop
    assert tuple(O.sxp_of(csf, s, h)) == sxp
    assert all(np.asarray(sxp) >= np.asarray(sx))  # the grid always covers the image


@pytest.mark.parametrize('geom', GEOMS, ids=str)
def test_partition_of_unity(geom):
    sx, csf, sf, ov = geom
    wa = O.Win2WinAux(sx, csf, ov)
    tot = O.ola(wa.ws, wa.csf, wa.sw, wa.nhop)
    np.testing.assert_allclose(tot, 1.0, rtol=0, atol=1e-12)


# ------------------------------------------------- index primitives (CUDA)
def _cuda_chop_pad(x, csf, sw, nhop, sz, offset):
    """gputools.cu:60-79 chop_pad_Kernel, thread by thread, after
    zeroPadKernel (gputools.cu:154-166) to sxp."""
    sxp = (nhop[0] * csf[0] + sw[0] - nhop[0], nhop[1] * csf[1] + sw[1] - nhop[1])
    dy = np.zeros(sxp)
    for i in range(sxp[0]):
        for j in range(sxp[1]):
            if i < x.shape[0] and j < x.shape[1]:
                dy[i, j] = x[i, j]
    xm, xn, xr = sz[0], sz[1], csf[0] * csf[1]
    wm, wn, om, on, cn = sw[0], sw[1], offset[0], offset[1], csf[1]
    dx = np.full(xr * xm * xn, np.nan)
    for i in range(xm * xr):
        for j in range(xn):
            k = i // xm
            m = k // cn
            n = k - m * cn
            if ((k * xm) + om - 1) < i < ((k * xm) + wm + om) and (on - 1) < j < (wn + on):
                dx[i * xn + j] = dy[(m * nhop[0]) - (k * xm) + i - om, (n * nhop[1]) + j - on]
            else:
                dx[i * xn + j] = 0.0
    return dx.reshape(xr, xm, xn)


def _cuda_ola(xs, csf, wm, nhop, offset):
    """gputools.cu:296-314 ola_Kernel_test, thread by thread."""
    outm = nhop[0] * csf[0] + wm[0] - nhop[0]
    outn = nhop[1] * csf[1] + wm[1] - nhop[1]
    inm, inn = xs.shape[1], xs.shape[2]
    flat = xs.ravel()
    out = np.zeros((outm, outn))
    for i in range(outm):
        for j in range(outn):
            ci = min(i // nhop[0], csf[0] - 1)
            cj = min(j // nhop[1], csf[1] - 1)
            res = 0.0
            d_i = 0
            while i < wm[0] + nhop[0] * (ci - d_i) and ci - d_i >= 0:
                d_j = 0
                while j < wm[1] + nhop[1] * (cj - d_j) and cj - d_j >= 0:
                    k = (ci - d_i) * csf[1] + (cj - d_j)
                    res += flat[k * inm * inn + (offset[0] + i + (d_i - ci) * nhop[0]) * inn
                                + offset[1] + j + (d_j - cj) * nhop[1]]
                    d_j += 1
                d_i += 1
            out[i, j] = res
    return out


@pytest.mark.parametrize('case', [((11, 13), (2, 3), (6, 7), (3, 4), (12, 16), (2, 3)),
                                  ((9, 9), (3, 3), (4, 4), (2, 2), (8, 8), (0, 0)),
                                  ((10, 7), (2, 2), (6, 5), (3, 3), (9, 10), (3, 1))])
def test_chop_pad_and_ola_bruteforce(case):
    sx, csf, sw, nhop, sz, off = case
    rng = np.random.default_rng(0)
    x = rng.standard_normal(sx)
    np.testing.assert_array_equal(O.chop_pad(x, csf, sw, nhop, sz, off), _cuda_chop_pad(x, csf, sw, nhop, sz, off))
    xs = rng.standard_normal((csf[0] * csf[1], sz[0], sz[1]))
    wm = (sz[0] - off[0], sz[1] - off[1])
    np.testing.assert_allclose(O.ola(xs, csf, wm, nhop, off), _cuda_ola(xs, csf, wm, nhop, off), rtol=1e-15, atol=1e-13)


def test_pad_stack_crop_stack_bruteforce():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 4, 5))
    sz, off = (9, 8), (2, 3)
    got = O.pad_stack(x, sz, off)
    # gputools.cu:184-201 pad_stack_Kernel
    pm, pn, xm, xn = sz[0], sz[1], 4, 5
    ref = np.zeros((3 * pm, pn))
    for i in range(pm * 3):
        for j in range(pn):
            k = i // pm
            if ((k * pm) + off[0] - 1) < i < ((k * pm) + xm + off[0]) and (off[1] - 1) < j < (xn + off[1]):
                ref[i, j] = x.ravel()[(i - off[0] + k * (xm - pm)) * xn + j - off[1]]
    np.testing.assert_array_equal(got.reshape(-1, pn), ref)
    back = O.crop_stack(got, (4, 5), off)
    np.testing.assert_array_equal(back, x)


def test_chop_ola_round_trip():
    """Chop with windows == 1 on their support, OLA with the same geometry and
    divide by the OLA of ones: the image comes back (gputools.py:1437-1449)."""
    rng = np.random.default_rng(2)
    sx, csf, sw, nhop = (40, 36), (3, 3), (20, 18), (10, 9)
    x = rng.random(sx)
    ch = O.chop_pad(x, csf, sw, nhop, sw)
    num = O.ola(ch, csf, sw, nhop)
    den = O.ola(np.ones_like(ch), csf, sw, nhop)
    np.testing.assert_allclose((num / den)[:sx[0], :sx[1]], x, rtol=1e-14)


# ------------------------------------------------------------- forward
@pytest.mark.parametrize('geom', GEOMS, ids=str)
@pytest.mark.parametrize('mode', ['valid', 'same', 'full'])
def test_invariant_psf_equals_conv2(geom, mode):
    """With one PSF for every patch the windows form a partition of unity and
    the operator is scipy's convolve2d (crop offset floor(sf/2) for 'same')."""
    sx, csf, sf, ov = geom
    wa = O.Win2WinAux(sx, csf, ov)
    rng = np.random.default_rng(3)
    x = rng.random(sx)
    f1 = rand_psfs(rng, 1, sf)[0]
    f = np.tile(f1, (csf[0] * csf[1], 1, 1))
    y = O.forward(x, f, wa, mode)
    full = ss.convolve2d(x, f1, 'full')
    if mode == 'valid':
        ref = ss.convolve2d(x, f1, 'valid')
    elif mode == 'full':
        ref = full
    else:
        o = (sf[0] // 2, sf[1] // 2)
        ref = full[o[0]:o[0] + sx[0], o[1]:o[1] + sx[1]]
        if sf[0] % 2 and sf[1] % 2:
            np.testing.assert_allclose(ref, ss.convolve2d(x, f1, 'same'), rtol=1e-13)
    assert rel_l2(y, ref) < 1e-13


def _direct_forward(x, f, wa, mode):
    """Brute force: sum over patches of convolve2d(window * patch, f_k),
    placed at the patch origin, cropped."""
    sf = f.shape[1:]
    sw = wa.sw
    sxp = wa.sxp
    xp = np.zeros(sxp)
    xp[:x.shape[0], :x.shape[1]] = x
    acc = np.zeros((sxp[0] + sf[0] - 1, sxp[1] + sf[1] - 1))
    for k, (p0, p1) in enumerate(O.patch_origins(wa.csf, wa.nhop)):
        patch = xp[p0:p0 + sw[0], p1:p1 + sw[1]] * wa.ws[k]
        acc[p0:p0 + sw[0] + sf[0] - 1, p1:p1 + sw[1] + sf[1] - 1] += ss.convolve2d(patch, f[k], 'full')
    sy = O.out_size(x.shape, sf, mode)
    oc = O.crop_offset(np.array(sf), mode)
    return acc[oc[0]:oc[0] + sy[0], oc[1]:oc[1] + sy[1]]


@pytest.mark.parametrize('geom', GEOMS[:5], ids=str)
@pytest.mark.parametrize('mode', ['valid', 'same', 'full'])
def test_forward_bruteforce(geom, mode):
    sx, csf, sf, ov = geom
    wa = O.Win2WinAux(sx, csf, ov)
    x, f, _ = case_inputs(4, sx, csf, sf, mode)
    assert rel_l2(O.forward(x, f, wa, mode), _direct_forward(x, f, wa, mode)) < 1e-13


def test_delta_psf_shift():
    sx, csf, sf = (64, 64), (2, 2), (7, 7)
    wa = O.Win2WinAux(sx, csf, 0.5)
    x = np.random.default_rng(5).random(sx)
    f = np.zeros((4,) + sf)
    f[:, 2, 5] = 1.0
    y = O.forward(x, f, wa, 'valid')
    # y[i,j] = x[i + sf0-1 - 2, j + sf1-1 - 5]
    np.testing.assert_allclose(y, x[4:4 + y.shape[0], 1:1 + y.shape[1]], rtol=0, atol=1e-13)


@pytest.mark.parametrize('geom', GEOMS, ids=str)
def test_Xf_equals_Fx(geom):
    sx, csf, sf, ov = geom
    wa = O.Win2WinAux(sx, csf, ov)
    x, f, _ = case_inputs(6, sx, csf, sf)
    assert rel_l2(O.OlaOracle(x, sf, 'valid', wa).cnv(f), O.OlaOracle(f, sx, 'valid', wa).cnv(x)) < 1e-13


def test_tile_independence():
    sx, csf, sf = (96, 96), (3, 3), (9, 9)
    wa = O.Win2WinAux(sx, csf, 0.5)
    x, f, g = case_inputs(7, sx, csf, sf)
    T = O.fft_tile(wa.sw, sf)
    for tile in [(T[0] * 2, T[1]), (57, 70), (T[0], T[1] * 4)]:
        assert rel_l2(O.forward(x, f, wa, tile=tile), O.forward(x, f, wa)) < 1e-13
        for exact in (True, False):
            assert rel_l2(O.adjoint_x(g, f, wa, sx, exact=exact, tile=tile),
                          O.adjoint_x(g, f, wa, sx, exact=exact)) < 1e-13
            assert rel_l2(O.adjoint_f(g, x, wa, sf, exact=exact, tile=tile),
                          O.adjoint_f(g, x, wa, sf, exact=exact)) < 1e-13


# ------------------------------------------------------------ transposes
@pytest.mark.parametrize('geom', GEOMS, ids=str)
@pytest.mark.parametrize('mode', ['valid', 'same'])
def test_exact_adjoint_dot_products(geom, mode):
    sx, csf, sf, ov = geom
    wa = O.Win2WinAux(sx, csf, ov)
    x, f, g = case_inputs(8, sx, csf, sf, mode)
    y = O.forward(x, f, wa, mode)
    lhs = np.vdot(y, g)
    assert abs(lhs - np.vdot(x, O.adjoint_x(g, f, wa, sx, mode, exact=True))) < 1e-12 * abs(lhs) + 1e-12
    assert abs(lhs - np.vdot(f, O.adjoint_f(g, x, wa, sf, mode, exact=True))) < 1e-12 * abs(lhs) + 1e-12


def _masked_correlations(g, wa, sf, mode, exact):
    """Residual patches G_k (full correlation support when exact, the
    reference's sw-sized chop at offset oc otherwise) by direct indexing."""
    oc = O.crop_offset(np.array(sf), mode)
    ext = np.asarray(wa.sw) + np.asarray(sf) - 1
    tiles = []
    for (p0, p1) in O.patch_origins(wa.csf, wa.nhop):
        G = np.zeros(tuple(ext))
        for b0 in range(ext[0]):
            for b1 in range(ext[1]):
                if not exact and not (oc[0] <= b0 < oc[0] + wa.sw[0] and oc[1] <= b1 < oc[1] + wa.sw[1]):
                    continue
                i, j = p0 + b0 - oc[0], p1 + b1 - oc[1]
                if 0 <= i < g.shape[0] and 0 <= j < g.shape[1]:
                    G[b0, b1] = g[i, j]
        tiles.append(G)
    return tiles


@pytest.mark.parametrize('geom', GEOMS[:4], ids=str)
@pytest.mark.parametrize('mode', ['valid', 'same'])
@pytest.mark.parametrize('exact', [True, False])
def test_adjoints_bruteforce_correlate2d(geom, mode, exact):
    """fbar_k = correlate2d(G_k, xw_k, 'valid'); xbar = OLA(ws_k *
    correlate2d(G_k, f_k, 'valid')) cropped to the image."""
    sx, csf, sf, ov = geom
    wa = O.Win2WinAux(sx, csf, ov)
    x, f, g = case_inputs(9, sx, csf, sf, mode)
    tiles = _masked_correlations(g, wa, sf, mode, exact)
    xp = np.zeros(wa.sxp)
    xp[:sx[0], :sx[1]] = x
    fb = []
    xb = np.zeros(wa.sxp)
    for k, (p0, p1) in enumerate(O.patch_origins(wa.csf, wa.nhop)):
        xw = xp[p0:p0 + wa.sw[0], p1:p1 + wa.sw[1]] * wa.ws[k]
        fb.append(ss.correlate2d(tiles[k], xw, 'valid'))
        xb[p0:p0 + wa.sw[0], p1:p1 + wa.sw[1]] += wa.ws[k] * ss.correlate2d(tiles[k], f[k], 'valid')
    assert rel_l2(O.adjoint_f(g, x, wa, sf, mode, exact=exact), np.stack(fb)) < 1e-12
    assert rel_l2(O.adjoint_x(g, f, wa, sx, mode, exact=exact), xb[:sx[0], :sx[1]]) < 1e-12


def test_finite_difference_gradcheck():
    """loss = 1/2 ||B(x,f) - y||^2 on a 96x96 image, 3x3 patches, 9x9 PSFs:
    the exact transposes of the residual are the gradients (central
    differences, FP64)."""
    sx, csf, sf = (96, 96), (3, 3), (9, 9)
    wa = O.Win2WinAux(sx, csf, 0.5)
    x, f, _ = case_inputs(10, sx, csf, sf)
    yobs = np.random.default_rng(11).random(O.out_size(sx, sf, 'valid'))

    def loss(xx, ff):
        r = O.forward(xx, ff, wa) - yobs
        return 0.5 * np.vdot(r, r)

    r = O.forward(x, f, wa) - yobs
    gx = O.adjoint_x(r, f, wa, sx, exact=True)
    gf = O.adjoint_f(r, x, wa, sf, exact=True)
    rng = np.random.default_rng(12)
    h = 1e-5
    for _ in range(6):
        i, j = rng.integers(0, sx[0]), rng.integers(0, sx[1])
        e = np.zeros(sx)
        e[i, j] = h
        fd = (loss(x + e, f) - loss(x - e, f)) / (2 * h)
        assert abs(fd - gx[i, j]) <= 1e-6 * (abs(gx[i, j]) + 1e-3)
        k, a, b = rng.integers(0, 9), rng.integers(0, sf[0]), rng.integers(0, sf[1])
        e = np.zeros_like(f)
        e[k, a, b] = h
        fd = (loss(x, f + e) - loss(x, f - e)) / (2 * h)
        assert abs(fd - gf[k, a, b]) <= 1e-6 * (abs(gf[k, a, b]) + 1e-3)


def test_reference_errors():
    wa = O.Win2WinAux((64, 64), (2, 2), 0.5)
    with pytest.raises(IOError):
        O.OlaOracle(np.ones((3, 5, 5)), (64, 64), 'valid', wa)   # olaGPU.py:110-111
    with pytest.raises(NotImplementedError):
        O.OlaOracle(np.ones((4, 5, 5)), (64, 64), 'bogus', wa)   # olaGPU.py:136
    with pytest.raises(IOError):
        O.OlaOracle(np.ones((4, 5, 5)), (60, 64), 'valid', wa)   # olaGPU.py:124-125
    # olaGPU.py:285-286 stops cnvtp at valid/same; full extends the same rule
    F = O.OlaOracle(np.ones((4, 5, 5)) / 25, (64, 64), 'full', wa)
    assert F.cnvtp(np.ones((68, 68))).shape == (64, 64)


def test_full_mode_transposes_dot_product():
    """Full-mode exact transposes satisfy <Fx, g> = <x, F'g> = <f, X'g>."""
    sx, csf, sf = (40, 48), (2, 3), (5, 7)
    wa = O.Win2WinAux(sx, csf, 0.5)
    rng = np.random.default_rng(17)
    x = rng.random(sx)
    f = rng.random((6,) + sf)
    y = O.forward(x, f, wa, 'full')
    g = rng.standard_normal(y.shape)
    lhs = np.sum(y * g)
    xb = O.adjoint_x(g, f, wa, sx, 'full', exact=True)
    fb = O.adjoint_f(g, x, wa, sf, 'full', exact=True)
    assert abs(lhs - np.sum(x * xb)) < 1e-12 * abs(lhs)
    assert abs(lhs - np.sum(f * fb)) < 1e-12 * abs(lhs)


def test_ola_flat_matches_kernel_loops():
    """ola_flat against a literal transcription of the loops of
    gputools.cu:295-314 (covering patch ci - d_i while i < wm + hop (ci - d_i),
    flat read k*inm*inn + r*inn + c, reads past the stack give 0), with an
    extent wm larger than the slices (the F.cnvtp call, olaGPU.py:307-309)."""
    rng = np.random.default_rng(21)
    csf, sw, hop = (2, 3), (6, 6), (3, 3)
    xs = rng.random((6,) + sw)
    wm = (8, 8)
    flat = xs.ravel()
    sx = O.sxp_of(csf, wm, hop)
    ref = np.zeros((int(sx[0]), int(sx[1])))
    for i in range(ref.shape[0]):
        for j in range(ref.shape[1]):
            ci = min(i // hop[0], csf[0] - 1)
            cj = min(j // hop[1], csf[1] - 1)
            d_i = 0
            while i < wm[0] + hop[0] * (ci - d_i) and ci - d_i >= 0:
                d_j = 0
                while j < wm[1] + hop[1] * (cj - d_j) and cj - d_j >= 0:
                    r = i + (d_i - ci) * hop[0]
                    c = j + (d_j - cj) * hop[1]
                    k = (ci - d_i) * csf[1] + (cj - d_j)
                    e = k * sw[0] * sw[1] + r * sw[1] + c
                    ref[i, j] += flat[e] if e < flat.size else 0.0
                    d_j += 1
                d_i += 1
    np.testing.assert_allclose(O.ola_flat(xs, csf, wm, hop), ref, rtol=1e-14, atol=1e-14)
