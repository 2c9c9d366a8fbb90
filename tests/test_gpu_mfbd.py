"""GPU parity of the online MFBD loop (vmbd_chimney.py:110-147) and of the
device PBB solver against the FP64 oracles.

Iterated solvers amplify rounding: with every operator output rounded to
float32, the FP64 oracle's PBB iterates move by about 1e-5 (10-20
iterations) to 1e-4 (40 iterations) in relative L2, so these comparisons use
1e-3 after at most 10 iterations."""
import numpy as np
import pytest
import scipy.optimize as so

from oracle import mfbd_oracle as M
from oracle import ola_oracle as O
from tests.helpers import case_inputs, rel_l2

pytestmark = pytest.mark.gpu

TOL_ITER = 1e-3


def _t(a, cuda):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=cuda)


def _n(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _scene(seed, sy, sf, csf):
    """Latent x (sy + sf - 1), PSFs, observation y with a saturated strip, and
    a blurred first estimate y0."""
    sx = (sy[0] + sf[0] - 1, sy[1] + sf[1] - 1)
    x, f, _ = case_inputs(seed, sx, csf, sf)
    wa = O.Win2WinAux(sx, csf, 0.5)
    y = O.forward(x, f, wa) * 1.6 + 0.01 * np.random.default_rng(seed + 1).standard_normal(sy)
    # saturated pixels (masked out), narrower than the PSF so that every
    # latent pixel still sees unmasked samples: where all samples a pixel
    # sees are masked, F'(mask y) and F'(mask F x) are both zero and their
    # ratio is float32 FFT round-off over round-off (arbitrary, also in the
    # reference), so such a scene has no well-defined answer to compare
    y[28:31, 10:50] = np.maximum(y[28:31, 10:50], 1.2)
    y0 = O.forward(x, np.full_like(f, 1.0 / (sf[0] * sf[1])), wa)
    return sx, wa, y.astype(np.float32).astype(np.float64), y0.astype(np.float32).astype(np.float64)


def _rl_rel(x_dev, x_ref, x0, f, y, wa, sf, exact):
    """Relative L2 of the multiplicative x-step over the pixels it determines.

    A latent pixel that no unmasked observation reaches with nonzero PSF
    weight has F'(mask y) = F'(mask F x) = 0 exactly; its update is float32
    round-off over round-off (0/0, in the reference as well), so such pixels
    (den below 1e-6 of its maximum) are left out; they must be rare."""
    F = O.OlaOracle(f, x0.shape, 'valid', wa)
    mask = (np.asarray(y) < 1).astype(np.float64)
    den = F.cnvtp(mask * F.cnv(x0), exact=exact)
    ok = den > 1e-6 * den.max()
    assert ok.mean() > 0.99
    return rel_l2(x_dev[ok], x_ref[ok])


@pytest.mark.parametrize('exact,f_init', [(False, 'xty'), (True, 'xty'), (False, 'warm')])
def test_frame_pbb_rl_vs_oracle(cuda, exact, f_init):
    from vmbd_amd.mfbd import OnlineMFBD
    sy, sf, csf = (64, 64), (7, 7), (3, 3)
    print(f'case mfbd pbb+rl exact={exact} f_init={f_init}', flush=True)
    sx, wa, y, y0 = _scene(11, sy, sf, csf)
    m = OnlineMFBD(y0, sf, csf, f_solver='pbb', x_solver='rl', f_iters=10, f_beta=0.1, exact=exact, f_init=f_init)
    x0 = _n(m.x)
    np.testing.assert_array_equal(x0, np.pad(y0.astype(np.float32), ((3, 3), (3, 3)), mode='edge'))
    f0 = None if f_init == 'xty' else np.full((9,) + sf, 1.0 / 49, dtype=np.float32).astype(np.float64)
    rng = np.random.default_rng(15)
    for frame in range(2 if f_init == 'warm' else 1):
        yi = y if frame == 0 else y + 0.01 * rng.standard_normal(y.shape)
        yi = yi.astype(np.float32).astype(np.float64)
        x1, f1 = m.step(yi)
        fu = _n(m.fu)
        xr, fr = M.frame_step(x0, yi, wa, sf, fu, 10, 0.1, tol=1e-10, exact=exact, f0=f0)
        assert rel_l2(_n(f1), fr) < TOL_ITER, frame
        assert _rl_rel(_n(x1), xr, x0, fr, yi, wa, sf, exact) < TOL_ITER, frame
        # the x-step alone, from the PSFs the device produced: one pass, operator tolerance
        xr2 = M.rl_step(x0, _n(f1), yi, wa, sf, tol=1e-10, exact=exact)
        assert _rl_rel(_n(x1), xr2, x0, _n(f1), yi, wa, sf, exact) < 1e-5, frame
        x0, f0 = _n(x1), _n(f1)      # the next frame starts from the device state


def test_frame_pbb_pbb_vs_oracle(cuda):
    from vmbd_amd.mfbd import OnlineMFBD
    sy, sf, csf = (60, 72), (5, 5), (2, 3)
    print('case mfbd pbb+pbb', flush=True)
    sx, wa, y, y0 = _scene(12, sy, sf, csf)
    m = OnlineMFBD(y0, sf, csf, f_solver='pbb', x_solver='pbb', f_iters=8, x_iters=6, f_beta=0.05, exact=True)
    x0 = _n(m.x)
    x1, f1 = m.step(y)
    f0 = np.full((6,) + sf, np.float32(1.0 / 25)).astype(np.float64)
    xr, fr = M.frame_step(x0, y, wa, sf, _n(m.fu), 8, 0.05, exact=True, x_solver='pbb', xu=_n(m.xu), x_iters=6,
                          f0=f0)
    assert rel_l2(_n(f1), fr) < TOL_ITER
    assert rel_l2(_n(x1), xr) < TOL_ITER


def test_frame_with_prior_vs_oracle(cuda):
    """PBB f-step, PBB x-step and the bilateral + shock prior pass."""
    from vmbd_amd.mfbd import OnlineMFBD
    sy, sf, csf = (64, 64), (7, 7), (3, 3)
    print('case mfbd prior', flush=True)
    sx, wa, y, y0 = _scene(16, sy, sf, csf)
    m = OnlineMFBD(y0, sf, csf, f_solver='pbb', x_solver='pbb', f_iters=6, x_iters=4, f_beta=0.1, exact=True,
                   prior='bilateral_shock')
    x0 = _n(m.x)
    x1, f1 = m.step(y)
    f0 = np.full((9,) + sf, np.float32(1.0 / 49)).astype(np.float64)
    xr, fr = M.frame_step(x0, y, wa, sf, _n(m.fu), 6, 0.1, exact=True, x_solver='pbb', xu=_n(m.xu), x_iters=4,
                          f0=f0, prior=True)
    assert rel_l2(_n(f1), fr) < TOL_ITER
    assert rel_l2(_n(x1), xr) < TOL_ITER


def test_capture_replays_match_eager(cuda):
    """Three frames replayed from one recorded graph equal three eager frames
    bit for bit (deterministic reductions, same random draws)."""
    import torch
    from vmbd_amd.mfbd import OnlineMFBD
    sy, sf, csf = (64, 64), (7, 7), (3, 3)
    print('case mfbd capture', flush=True)
    _, _, y, y0 = _scene(13, sy, sf, csf)
    kw = dict(f_solver='pbb', x_solver='rl', f_iters=6, f_beta=0.1, seed=7)
    a = OnlineMFBD(y0, sf, csf, capture=False, **kw)
    b = OnlineMFBD(y0, sf, csf, capture=True, **kw)
    rng = np.random.default_rng(5)
    for i in range(3):
        yi = y + 0.01 * rng.standard_normal(y.shape)
        xa, fa = a.step(yi)
        xa, fa = xa.clone(), fa.clone()
        torch.cuda.synchronize()
        xb, fb = b.step(yi)
        torch.cuda.synchronize()
        assert torch.equal(xa, xb), i
        assert torch.equal(fa, fb), i


def test_frame_lbfgsb(cuda):
    """The reference's f-solver (host L-BFGS-B over device operators): the
    PSFs come back normalised and nonnegative, and the x-step from them
    matches the oracle's multiplicative step."""
    from vmbd_amd.mfbd import OnlineMFBD
    sy, sf, csf = (64, 64), (7, 7), (3, 3)
    print('case mfbd lbfgsb', flush=True)
    sx, wa, y, y0 = _scene(14, sy, sf, csf)
    m = OnlineMFBD(y0, sf, csf, f_solver='lbfgsb', x_solver='rl', optiter=30, f_beta=0.1)
    x0 = _n(m.x)
    x1, f1 = m.step(y)
    f1 = _n(f1)
    assert np.all(f1 >= 0)
    np.testing.assert_allclose(f1.sum(axis=(1, 2)), 1.0, rtol=1e-5)
    xr = M.rl_step(x0, f1, y, wa, sf, tol=1e-10)
    assert _rl_rel(_n(x1), xr, x0, f1, y, wa, sf, False) < 1e-5


def test_pbb_class_converges_to_nnls(cuda):
    """optGPU.PBB on the dense MVM objective (optGPU.py:19-31) reaches the
    NNLS optimum (scaled so the 1e-5 BB guard sits below float32 accuracy)."""
    import torch
    from vmbd_amd import optGPU
    print('case PBB class', flush=True)
    rng = np.random.default_rng(1)
    A = 100 * rng.standard_normal((40, 12))
    b = 100 * rng.standard_normal(40)
    ref, _ = so.nnls(A, b)
    opt = optGPU.Solopt()
    opt.maxiter = 300
    opt.tolo = 0.0
    obj = optGPU.MVM_Objective(_t(A, cuda), _t(b, cuda))
    res = optGPU.PBB(obj, _t(rng.random(12) + 0.1, cuda), opt)
    assert res.iter == 300
    assert rel_l2(_n(res.result), ref) < 1e-4
    assert torch.all(res.result >= 0)
