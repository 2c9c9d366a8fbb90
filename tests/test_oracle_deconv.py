"""Pins of the closed-form deconvolution oracle (oracle/deconv_oracle.py) and
of the L-BFGS-B objective/gradient oracle, by checks that do not reuse the
oracle's own arithmetic:

* every Fourier division is the normal equation of a per-patch Tikhonov
  least-squares problem over the whole FFT tile.  Here that problem is built
  as dense circulant matrices, entry by entry from the filter taps, and
  solved with numpy.linalg (no FFT anywhere);
* the chop / window / overlap-add composition around the solves is
  restated with explicit per-pixel loops;
* a csf = (1, 1) grid, whose single normalised window is 1 everywhere, turns
  xdirect into one plain Tikhonov deconvolution of the whole image;
* the objective's gradient is checked against FP64 central differences of
  the objective, which pins the kept sign quirk of the reference's
  gradient (olaGPU.py:808-811) as an identity.

Reference: olaGPU.py:416-733, 756-811; gputools.cu:316-406.
"""
import numpy as np
import pytest

from oracle import deconv_oracle as DO
from oracle import filters_oracle as FO
from oracle import ola_oracle as O


# ------------------------------------------------------------ dense helpers
def circ(taps, T):
    """Matrix of z -> k (*) z, the circular convolution on a T0 x T1 tile of
    the kernel k = taps placed at the tile origin; entry ((i,j),(a,b)) =
    k[(i-a) mod T0, (j-b) mod T1]."""
    T0, T1 = int(T[0]), int(T[1])
    k = np.zeros((T0, T1))
    taps = np.asarray(taps, dtype=np.float64)
    k[:taps.shape[0], :taps.shape[1]] = taps
    i, j, a, b = np.meshgrid(np.arange(T0), np.arange(T1), np.arange(T0), np.arange(T1), indexing='ij')
    return k[(i - a) % T0, (j - b) % T1].reshape(T0 * T1, T0 * T1)


def chop_loops(img, csf, sw, nhop, T, off=(0, 0)):
    """Patch k = (m, n) of img (zero outside the image), rows and columns
    [0, sw), placed at tile position off + (a, b): explicit loops."""
    img = np.asarray(img, dtype=np.float64)
    H, W = img.shape
    P = csf[0] * csf[1]
    out = np.zeros((P, T[0], T[1]))
    for m in range(csf[0]):
        for n in range(csf[1]):
            k = m * csf[1] + n
            for a in range(int(sw[0])):
                for b in range(int(sw[1])):
                    r, c = off[0] + a, off[1] + b
                    i, j = m * int(nhop[0]) + a, n * int(nhop[1]) + b
                    if r < T[0] and c < T[1] and i < H and j < W:
                        out[k, r, c] = img[i, j]
    return out


def ola_loops(zs, csf, sw, nhop, sx):
    """Sum of patch k's (sw) block at its origin, cropped to sx: loops."""
    out = np.zeros((int(sx[0]), int(sx[1])))
    for m in range(csf[0]):
        for n in range(csf[1]):
            k = m * csf[1] + n
            for a in range(int(sw[0])):
                for b in range(int(sw[1])):
                    i, j = m * int(nhop[0]) + a, n * int(nhop[1]) + b
                    if i < sx[0] and j < sx[1]:
                        out[i, j] += zs[k, a, b]
    return out


def tikhonov(mats, rhs_tiles, reg, beta):
    """argmin_z sum_i |A_i z - y_i|^2 + z' reg z + beta |z|^2 (dense)."""
    n = mats[0].shape[1]
    lhs = sum(A.T @ A for A in mats) + reg + beta * np.eye(n)
    rhs = sum(A.T @ y.ravel() for A, y in zip(mats, rhs_tiles))
    return np.linalg.solve(lhs, rhs)


def grad_reg(filters, T, alpha):
    if alpha == 0:
        return 0.0
    return alpha * sum(circ(f, T).T @ circ(f, T) for f in filters)


def _inputs(seed, sx, csf, sf):
    rng = np.random.default_rng(seed)
    x = rng.random(sx)
    f = rng.random((csf[0] * csf[1],) + sf)
    f /= f.sum(axis=(1, 2), keepdims=True)
    return rng, x, f


DX, DY = [[-1.0, 1.0]], [[-1.0], [1.0]]


# ------------------------------------------------------------------ xdirect
@pytest.mark.parametrize('alpha,beta', [(0.0, 1e-3), (0.05, 0.01)])
def test_xdirect_single_patch_is_tikhonov_deconvolution(alpha, beta):
    """csf = (1, 1): the window is 1 on the whole patch, so 'xdirect' of an F
    instance is one Tikhonov deconvolution of y (placed at the valid offset
    sf - 1) by f over the 16 x 16 tile."""
    sx, csf, sf = (10, 10), (1, 1), (5, 5)
    wa = O.Win2WinAux(sx, csf, 0.5)
    assert tuple(wa.sw) == sx and np.array_equal(wa.ws, np.ones((1,) + sx))
    rng, x, f = _inputs(1, sx, csf, sf)
    Z = O.OlaOracle(f, sx, 'valid', wa)
    T = Z.T
    assert T == (16, 16)
    y = rng.random(O.out_size(sx, sf, 'valid'))
    got = DO.xdirect(Z, y, alpha, beta)
    ytile = np.zeros(T)
    ytile[4:10, 4:10] = y
    z = tikhonov([circ(f[0], T)], [ytile], grad_reg([DX, DY], T, alpha), beta).reshape(T)
    np.testing.assert_allclose(got, z[:10, :10], rtol=1e-9, atol=1e-12)


def test_xdirect_patch_grid_f_and_x_instances():
    """2 x 2 patches: per patch, the residual patch (at the valid offset)
    times the window placed at tile offset 0 (the reference's placement,
    olaGPU.py:577-581), a dense Tikhonov solve against the cached operand,
    then crop to sw and overlap-add (F instance) or crop to sf (X instance)."""
    sx, csf, sf = (18, 18), (2, 2), (5, 5)
    alpha, beta = 0.05, 0.01
    wa = O.Win2WinAux(sx, csf, 0.5)
    sw, nhop = tuple(wa.sw), tuple(wa.nhop)
    rng, x, f = _inputs(2, sx, csf, sf)
    y = rng.random(O.out_size(sx, sf, 'valid'))
    F = O.OlaOracle(f, sx, 'valid', wa)
    T = F.T
    assert T == (16, 16)
    reg = grad_reg([DX, DY], T, alpha)
    wt = np.zeros((4,) + T)
    wt[:, :sw[0], :sw[1]] = wa.ws
    rhs = chop_loops(y, csf, sw, nhop, T, (4, 4)) * wt
    zs = np.stack([tikhonov([circ(f[k], T)], [rhs[k]], reg, beta).reshape(T)[:sw[0], :sw[1]] for k in range(4)])
    np.testing.assert_allclose(DO.xdirect(F, y, alpha, beta), ola_loops(zs, csf, sw, nhop, sx), rtol=1e-9,
                               atol=1e-12)
    # X instance: operand = windowed image patches at offset 0, result cropped to sf
    X = O.OlaOracle(x, sf, 'valid', wa)
    xt = chop_loops(x, csf, sw, nhop, T) * wt
    fs = np.stack([tikhonov([circ(xt[k], T)], [rhs[k]], reg, beta).reshape(T)[:sf[0], :sf[1]] for k in range(4)])
    np.testing.assert_allclose(DO.xdirect(X, y, alpha, beta), fs, rtol=1e-9, atol=1e-12)


# ------------------------------------------------------------------ gdirect
@pytest.mark.parametrize('alpha', [0.0, 0.03])
def test_gdirect_dense(alpha):
    """PSF from gradient images: per patch, the two gradient operands of x
    (offset 0) against the gradients of the tapered y (valid offset),
    normal equations of both data terms plus the 3x2 / 2x3 filter energy and
    beta, crop to sf (olaGPU.py:416-511)."""
    sx, csf, sf = (18, 18), (2, 2), (5, 5)
    beta = 0.02
    wa = O.Win2WinAux(sx, csf, 0.5)
    sw, nhop = tuple(wa.sw), tuple(wa.nhop)
    rng, x, f = _inputs(4, sx, csf, sf)
    y = rng.random(O.out_size(sx, sf, 'valid'))
    X = O.OlaOracle(x, sf, 'valid', wa)
    T = X.T
    yt = FO.edgetaper(y, 2 * np.asarray(sf))
    xx, xy = np.diff(x, axis=1) * -np.sqrt(0.5), np.diff(x, axis=0) * -np.sqrt(0.5)
    yx, yy = np.diff(yt, axis=1) * -np.sqrt(0.5), np.diff(yt, axis=0) * -np.sqrt(0.5)
    Ax = chop_loops(xx, csf, sw, nhop, T)
    Ay = chop_loops(xy, csf, sw, nhop, T)
    Bx = chop_loops(yx, csf, sw, nhop, T, (4, 4))
    By = chop_loops(yy, csf, sw, nhop, T, (4, 4))
    reg = grad_reg([[[-1, 1], [-1, 1], [-1, 1]], [[-1, -1, -1], [1, 1, 1]]], T, alpha)
    ref = np.stack([tikhonov([circ(Ax[k], T), circ(Ay[k], T)], [Bx[k], By[k]], reg, beta).reshape(T)
                    [:sf[0], :sf[1]] for k in range(4)])
    np.testing.assert_allclose(DO.gdirect(X, y, alpha, beta), ref, rtol=1e-9, atol=1e-12)
    with pytest.raises(Exception, match='direct mode'):
        DO.gdirect(O.OlaOracle(f, sx, 'valid', wa), y, alpha, beta)


# ------------------------------------------------------------------- direct
def test_direct_crop_and_replicate_pad():
    """'direct' = max(xdirect(y), 0) / clip(taper(xdirect(taper(1))), .01, 10),
    cropped by floor((factor sf - 1)/2) to sz - factor sf - 1 and padded back
    by edge replication with floor(pad/2) before (olaGPU.py:515-546), the
    crop and the padding as explicit loops."""
    sx, csf, sf = (18, 18), (2, 2), (5, 5)
    wa = O.Win2WinAux(sx, csf, 0.5)
    rng, x, f = _inputs(5, sx, csf, sf)
    F = O.OlaOracle(f, sx, 'valid', wa)
    y = rng.random(O.out_size(sx, sf, 'valid'))
    for factor in (1, 2):
        got = DO.direct(F, y, 0.05, 0.01, factor=factor)
        yd = np.maximum(DO.xdirect(F, y, 0.05, 0.01), 0)
        cd = DO.xdirect(F, FO.edgetaper(np.ones(y.shape), 2 * np.asarray(sf)), 0.05, 0.01)
        cd = np.clip(FO.edgetaper(cd, 2 * np.asarray(sf)), 0.01, 10.0)
        q = yd / cd
        H, W = q.shape
        c0 = (factor * sf[0] - 1) // 2
        h = H - factor * sf[0] - 1
        w = W - factor * sf[1] - 1
        ph, pw = H - h, W - w
        ref = np.zeros((H, W))
        for i in range(H):
            for j in range(W):
                ii = min(max(i - ph // 2, 0), h - 1)
                jj = min(max(j - pw // 2, 0), w - 1)
                ref[i, j] = q[c0 + ii, c0 + jj]
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15)


# ------------------------------------------------------------------- sparse
def _sparse_dense(F, f, y, alpha, gamma, csf, sw, nhop, ws, sx, sf):
    """The half-quadratic loop of olaGPU.py:613-733 with every Fourier
    division replaced by a dense solve and every chop / OLA by loops."""
    T = F.T
    P = csf[0] * csf[1]
    r = np.sqrt(0.5)
    Gx, Gy = circ(r * np.array(DX), T), circ(r * np.array(DY), T)
    C = [circ(f[k], T) for k in range(P)]
    reg = Gx.T @ Gx + Gy.T @ Gy
    y = FO.edgetaper(y, 2 * np.asarray(sf))
    const = FO.edgetaper(np.ones(y.shape), 2 * np.asarray(sf))
    off = (sf[0] - 1, sf[1] - 1)
    wt = np.zeros((P,) + T)
    wt[:, off[0]:off[0] + sw[0], off[1]:off[1] + sw[1]] = ws[:, :T[0] - off[0], :T[1] - off[1]]
    yp = chop_loops(y, csf, sw, nhop, T, off) * wt
    cp = chop_loops(const, csf, sw, nhop, T, off) * wt
    cs = np.stack([tikhonov([C[k]], [cp[k]], alpha * reg, gamma).reshape(T)[:sw[0], :sw[1]]
                   for k in range(P)]) * ws
    const = ola_loops(cs, csf, sw, nhop, sx)
    const = 0.01 / np.maximum(const, 0.01)
    x = y.copy()
    beta = 1.0
    n_iter = 0
    while beta < 2.0 ** 8:
        gx = FO.shrink23(-np.sqrt(0.5) * np.diff(x, axis=1), beta)
        gy = FO.shrink23(-np.sqrt(0.5) * np.diff(x, axis=0), beta)
        wx = chop_loops(gx, csf, sw, nhop, T)
        wy = chop_loops(gy, csf, sw, nhop, T)
        ab = alpha * beta
        zs = []
        for k in range(P):
            lhs = C[k].T @ C[k] / ab + reg + gamma * np.eye(C[k].shape[0])
            rhs = Gx.T @ wx[k].ravel() + Gy.T @ wy[k].ravel() + C[k].T @ yp[k].ravel() / ab
            zs.append(np.linalg.solve(lhs, rhs).reshape(T)[:sw[0], :sw[1]])
        x = np.maximum(ola_loops(np.stack(zs) * ws, csf, sw, nhop, sx) * const, 0.0)
        beta *= 2.0 * np.sqrt(2.0)
        n_iter += 1
    return x, n_iter


@pytest.mark.parametrize('sx,csf', [((10, 10), (1, 1)), ((18, 18), (2, 2))])
def test_sparse_dense(sx, csf):
    """Half-quadratic splitting: six beta steps (1 -> x 2 sqrt 2 while below
    256), |x|^(2/3) w-steps, per-patch x-step as a dense solve, windows at
    the residual's offset, then OLA, the constant-image correction and the
    clip at zero."""
    sf = (5, 5)
    alpha, gamma = 0.002, 1e-4
    wa = O.Win2WinAux(sx, csf, 0.5)
    rng, x, f = _inputs(6, sx, csf, sf)
    F = O.OlaOracle(f, sx, 'valid', wa)
    y = O.forward(x, f, wa) + 0.01 * rng.standard_normal(O.out_size(sx, sf, 'valid'))
    ref, n = _sparse_dense(F, f, y, alpha, gamma, csf, tuple(wa.sw), tuple(wa.nhop), wa.ws, sx, sf)
    assert n == 6
    np.testing.assert_allclose(DO.sparse(F, y, alpha, gamma), ref, rtol=1e-8, atol=1e-11)


@pytest.mark.parametrize('alpha', [0.5, 2.0 / 3.0, 0.8])
def test_shrink_matches_minimiser(alpha):
    """kernel.cu:529-553 (alpha 0.5, 0.8) and kernel.cu:555-578 (alpha 2/3):
    the output minimises |x|^alpha + beta/2 (x - v)^2.  The nonzero local
    minimiser lies between 0 and v; it is found by bounded scalar
    minimisation to 1e-12.  A nonzero output must equal it to 1e-6 (four
    Newton steps from v converge that far; a wrong derivative does not) and
    must not lose to x = 0; a zero output must not lose to it either."""
    from scipy.optimize import minimize_scalar
    rng = np.random.default_rng(6)
    nonzero = zero = 0
    for beta in (1.0, 8.0, 64.0, 256.0):
        v = rng.uniform(-3, 3, 40)
        x = FO.shrink23(v, beta) if alpha == 2.0 / 3.0 else FO.shrink(v, beta, alpha)
        for vi, xi in zip(v, x):
            E = lambda t: abs(t) ** alpha + beta / 2 * (t - vi) ** 2
            r = minimize_scalar(E, bounds=(min(0.0, vi), max(0.0, vi)), method='bounded',
                                options={'xatol': 1e-12})
            tie = 1e-9 * (1 + E(0.0))
            if xi != 0:
                assert abs(xi - r.x) < 1e-6, (alpha, beta, vi, xi, r.x)
                assert E(r.x) <= E(0.0) + tie
                nonzero += 1
            else:
                assert E(0.0) <= E(r.x) + tie or abs(r.x) < 1e-6, (alpha, beta, vi, r.x)
                zero += 1
    assert nonzero > 60 and zero > 10


# -------------------------------------------------- objective and gradient
@pytest.mark.parametrize('inst', ['F', 'X'])
@pytest.mark.parametrize('beta', [0.0, 0.3])
def test_cnvinv_gradient_vs_central_differences(inst, beta):
    """With the true adjoint, gradfun(z) + 2 beta z equals the FP64 central
    difference gradient of objfun: the data term carries the true gradient's
    sign and the Tikhonov term the opposite one, as the reference returns it
    (olaGPU.py:808-811)."""
    sx, csf, sf = (18, 18), (2, 2), (3, 3)
    wa = O.Win2WinAux(sx, csf, 0.5)
    rng, x, f = _inputs(7, sx, csf, sf)
    y = rng.random(O.out_size(sx, sf, 'valid'))
    Z = O.OlaOracle(f, sx, 'valid', wa) if inst == 'F' else O.OlaOracle(x, sf, 'valid', wa)
    z = x if inst == 'F' else f
    g = DO.cnvinv_gradfun(Z, z, y, 0.0, beta, exact=True)
    h = 1e-5
    fd = np.zeros(z.size)
    for i in rng.choice(z.size, min(40, z.size), replace=False):
        e = np.zeros(z.size)
        e[i] = h
        e = e.reshape(z.shape)
        fd[i] = (DO.cnvinv_objfun(Z, z + e, y, 0.0, beta) - DO.cnvinv_objfun(Z, z - e, y, 0.0, beta)) / (2 * h)
        assert abs(g.ravel()[i] + 2 * beta * z.ravel()[i] - fd[i]) < 1e-7 * (1 + abs(fd[i])), i


def test_cnvinv_objective_closed_form_and_laplace_term():
    """objfun: 1/2 |y - Z z|^2 from the oracle forward, plus 1/2 alpha z.Lz
    with L applied as a dense matrix built from the 5-point stencil entry by
    entry ((4 c - sum of 4 neighbours)/sqrt(20) on interior pixels, rows of
    border pixels zero: the 'same' mode of kernel.cu:268-333), plus
    1/2 beta |z|^2; the X instance refuses alpha > 0 (NameError)."""
    sx, csf, sf = (18, 18), (2, 2), (3, 3)
    wa = O.Win2WinAux(sx, csf, 0.5)
    rng, x, f = _inputs(8, sx, csf, sf)
    y = rng.random(O.out_size(sx, sf, 'valid'))
    F = O.OlaOracle(f, sx, 'valid', wa)
    n0, n1 = sx
    Lm = np.zeros((n0 * n1, n0 * n1))
    for i in range(1, n0 - 1):
        for j in range(1, n1 - 1):
            Lm[i * n1 + j, i * n1 + j] = 4 / np.sqrt(20)
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                Lm[i * n1 + j, (i + di) * n1 + j + dj] = -1 / np.sqrt(20)
    r = y - F.cnv(x)
    want = 0.5 * np.sum(r * r) + 0.5 * 0.2 * x.ravel() @ Lm @ x.ravel() + 0.5 * 0.1 * np.sum(x * x)
    np.testing.assert_allclose(DO.cnvinv_objfun(F, x, y, 0.2, 0.1), want, rtol=1e-12)
    g = DO.cnvinv_gradfun(F, x, y, 0.2, 0.1)
    np.testing.assert_allclose(g, -(F.cnvtp(r) + 0.2 * (Lm @ x.ravel()).reshape(sx) + 0.1 * x), rtol=1e-12)
    with pytest.raises(NameError):
        DO.cnvinv_objfun(O.OlaOracle(x, sf, 'valid', wa), f, y, 0.2, 0.0)
