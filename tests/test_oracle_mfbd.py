"""Pins of the FP64 PBB and MFBD-step oracles (CPU only): closed-form BB
steps, convergence to scipy's NNLS optimum, and the multiplicative update's
fixed point."""
import numpy as np
import scipy.optimize as so

from oracle import mfbd_oracle as M
from oracle import ola_oracle as O
from tests.helpers import case_inputs


def test_pbb_isotropic_quadratic_one_step():
    """f(x) = c/2 |x - a|^2, a > 0: dg = c dx, so the BB2 step of the first
    update is 1/c up to the 1e-5 guard and lands on a."""
    rng = np.random.default_rng(0)
    a = rng.random(50) + 0.1
    c = 3.0
    u = rng.random(50)
    x = M.pbb_fixed(lambda z: c * (z - a), lambda z: 0.5 * c * np.sum((z - a) ** 2), rng.random(50), u, 1)
    np.testing.assert_allclose(x, a, rtol=1e-5)


def test_pbb_converges_to_nnls():
    """The 1e-5 guard in the BB denominators (optGPU.py:222-225) stops the
    iteration once sum dg^2 falls to its order, so the problem is scaled
    until that floor sits below the asserted accuracy."""
    rng = np.random.default_rng(1)
    A = 100 * rng.standard_normal((40, 12))
    b = 100 * rng.standard_normal(40)
    ref, _ = so.nnls(A, b)
    x = M.pbb_fixed(lambda z: A.T @ (A @ z - b), lambda z: 0.5 * np.sum((A @ z - b) ** 2), rng.random(12) + 0.1,
                    rng.random(12), 200)
    assert np.linalg.norm(x - ref) < 1e-6 * np.linalg.norm(ref)
    assert np.any(ref == 0)          # the bound is active somewhere: the projection is exercised


def test_f_step_matches_explicit_nnls():
    """The PBB f-step (before normalisation) solves the NNLS problem of the
    explicit matrix of f -> X f, assembled column by column."""
    sx, csf, sf = (24, 24), (2, 2), (3, 3)
    wa = O.Win2WinAux(sx, csf, 0.5)
    x, f, _ = case_inputs(2, sx, csf, sf)
    x = 100 * x                         # curvature well above the 1e-5 BB guard
    X = O.OlaOracle(x, sf, 'valid', wa)
    n = 4 * 9
    cols = []
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        cols.append(X.cnv(e.reshape(4, 3, 3)).ravel())
    A = np.stack(cols, 1)
    y = X.cnv(f) + 5 * np.random.default_rng(3).standard_normal(X.sy)
    beta = 50.0
    Ab = np.vstack([A, np.sqrt(beta) * np.eye(n)])
    bb = np.concatenate([y.ravel(), np.zeros(n)])
    ref, _ = so.nnls(Ab, bb)

    def obj(z):
        r = X.cnv(z) - y
        return 0.5 * np.sum(r * r) + 0.5 * beta * np.sum(z * z)

    z = M.pbb_fixed(lambda z: X.cnvtp(X.cnv(z) - y, exact=True) + beta * z, obj, X.cnvtp(y, exact=True),
                    np.random.default_rng(4).random((4, 3, 3)), 300)
    assert np.linalg.norm(z.ravel() - ref) < 5e-5 * np.linalg.norm(ref)
    assert np.any(ref == 0)


def test_rl_fixed_point():
    """Noise-free, unsaturated y = F x with the maximum of x in the interior:
    F'(y) = F'(F x), the factor is one and x is unchanged."""
    sx, csf, sf = (40, 40), (2, 2), (5, 5)
    wa = O.Win2WinAux(sx, csf, 0.5)
    x, f, _ = case_inputs(5, sx, csf, sf)
    x = 0.2 + 0.5 * x
    x[20, 20] = 0.95
    y = O.OlaOracle(f, sx, 'valid', wa).cnv(x)
    assert y.max() < 1
    xn = M.rl_step(x, f, y, wa, sf, tol=1e-16)
    np.testing.assert_allclose(xn, x, rtol=1e-11)


def test_rl_saturation_mask_and_clip():
    """Saturated pixels (y >= 1) drop out of both correlations; the update
    never exceeds the interior maximum."""
    sx, csf, sf = (40, 40), (2, 2), (5, 5)
    wa = O.Win2WinAux(sx, csf, 0.5)
    x, f, _ = case_inputs(6, sx, csf, sf)
    F = O.OlaOracle(f, sx, 'valid', wa)
    y = F.cnv(x) * 1.3
    y[:4] = 1.5
    xn = M.rl_step(x, f, y, wa, sf)
    assert xn.max() <= xn[5:-5, 5:-5].max() + 1e-15
    y2 = y.copy()
    y2[:4] = 7.0                       # other saturated values: same result
    np.testing.assert_array_equal(M.rl_step(x, f, y2, wa, sf), xn)


def test_chimney_oracle_trajectory_fixture():
    """The recorded FP64 trajectory of the reference's chimney loop (bars of
    the GPU chimney test): the generator reproduces its first frame, and
    the loop loses quality after its early peak (both transposes)."""
    import json
    import os
    from tests.golden import make_chimney_oracle_psnr as G
    rec = json.load(open(os.path.join(os.path.dirname(G.__file__), 'chimney_oracle_psnr.json')))
    for v in ('fixed', 'k1'):
        t = {int(k): x for k, x in rec[v].items()}
        assert t[100] < t[10] < t[5], v
    assert abs(G.trajectory('fixed', 1)['1'] - rec['fixed']['1']) <= 0.01
