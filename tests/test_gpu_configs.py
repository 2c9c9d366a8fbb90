"""GPU parity at the full sizes of the benchmark workloads, against the FP64
oracle on the same seeded inputs:

* streaming MFBD geometry: one 2048^2 frame, 16 x 16 patches, 63 x 63 PSFs,
  512 x 512 tiles (the multi-round streamed overlap-add: 24 covering patch
  pairs per output row, three rounds of eight transforms), forward, both
  exact adjoints, both cached forms, and its bilateral + shock prior pass;
* the 64-frame mixed batch's six geometries (512^2 / 768 x 1024 / 1024^2,
  6 x 6 .. 12 x 12 patches, 15 / 25 / 41 PSFs, including the odd window
  widths of 768 x 1024 and its 256 x 512 tiles) in one MixedBatch with three
  streams, replayed from its graph;
* one online frame at 512^2 (8 x 8 patches, 31 x 31 PSFs, 20 + 20 projected
  BB iterations), and the plain per-operator form on that geometry.

Operator outputs: relative L2 <= 1e-5.  Iterated solvers: the tolerance is
measured, not assumed: the oracle is run a second time with its FFTs and
spectral products in float32 (ola_oracle.set_fft(single=True)), and the
GPU result must lie within ten times that run's distance from the FP64
answer (and never needs to be closer than 1e-5).
"""
import os

import numpy as np
import pytest

from oracle import filters_oracle as FO
from oracle import mfbd_oracle as M
from oracle import ola_oracle as O
from tests.helpers import rel_l2

# the FP64 oracle takes tens of seconds at these sizes
pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

TOL = 1e-5
WORKERS = min(16, os.cpu_count() or 1)


@pytest.fixture(autouse=True)
def _threads():
    old = O.set_fft(WORKERS)
    yield
    O.set_fft(*old)


def _t(a, cuda):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=cuda)


def _n(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _f32(a):
    return np.float32(a).astype(np.float64)


def _scene(seed, sx, csf, sf, walk):
    from vmbd_amd import synth
    rng = np.random.default_rng(seed)
    x = synth.piecewise_smooth(sx, rng)
    f = synth.motion_psfs(csf[0] * csf[1], sf, rng, steps=walk)
    return rng, _f32(x), _f32(f)


def test_streamed_tile512_frame(cuda):
    """2048^2, 16 x 16 patches, 63^2 PSFs, tile 512: every operator entry
    point against the oracle."""
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    sx, csf, sf = (2048, 2048), (16, 16), (63, 63)
    print('case 2048^2 16x16 sf63 tile512', flush=True)
    wa = win2winaux(sx, csf, 0.5)
    plan = OlaPlan(sx, sf, wa, 'valid', device=cuda)
    assert plan.tile == (512, 512)
    owa = O.Win2WinAux(sx, csf, 0.5)
    rng, x, f = _scene(51, sx, csf, sf, 60)
    g = _f32(0.01 * rng.standard_normal(plan.sy))
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
    y = plan.forward(xt, ft)
    xb, fb = plan.adjoint_xf(gt, xt, ft, exact=True)
    Fs, Xs = plan.spectra('F', ft), plan.spectra('X', xt)
    yF, yX = plan.forward_cached('F', Fs, xt), plan.forward_cached('X', Xs, ft)
    xbF = plan.adjoint_cached('F', Fs, gt, exact=True)
    fbX = plan.adjoint_cached('X', Xs, gt, exact=True)
    xbL = plan.adjoint_cached('F', Fs, gt, exact=False)
    y_ref = O.forward(x, f, owa)
    xb_ref = O.adjoint_x(g, f, owa, sx, exact=True)
    fb_ref = O.adjoint_f(g, x, owa, sf, exact=True)
    errs = dict(y=rel_l2(_n(y), y_ref), yF=rel_l2(_n(yF), y_ref), yX=rel_l2(_n(yX), y_ref),
                xb=rel_l2(_n(xb), xb_ref), xbF=rel_l2(_n(xbF), xb_ref), fb=rel_l2(_n(fb), fb_ref),
                fbX=rel_l2(_n(fbX), fb_ref),
                xb_local=rel_l2(_n(xbL), O.adjoint_x(g, f, owa, sx, exact=False)))
    print('rel L2', {k: f'{v:.2e}' for k, v in errs.items()}, flush=True)
    assert all(v <= TOL for v in errs.values()), errs


def test_streamed_rounds_tile512(cuda):
    """The multi-round geometry of the OLA tests (60 covering patch pairs per
    row) on 512 x 512 tiles: eight transforms per round, eight rounds."""
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    sx, csf, sf = (200, 400), (4, 24), (9, 9)
    print('case streamed rounds tile 512', flush=True)
    wa = win2winaux(sx, csf, 0.75)
    plan = OlaPlan(sx, sf, wa, 'valid', tile=(512, 512), device=cuda)
    owa = O.Win2WinAux(sx, csf, 0.75)
    rng = np.random.default_rng(52)
    x, g = _f32(rng.random(sx)), _f32(rng.standard_normal(plan.sy))
    f = rng.random((plan.P,) + sf)
    f = _f32(f / f.sum(axis=(1, 2), keepdims=True))
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
    assert rel_l2(_n(plan.forward(xt, ft)), O.forward(x, f, owa)) <= TOL
    xb, fb = plan.adjoint_xf(gt, xt, ft, exact=True)
    assert rel_l2(_n(xb), O.adjoint_x(g, f, owa, sx, exact=True)) <= TOL
    assert rel_l2(_n(fb), O.adjoint_f(g, x, owa, sf, exact=True)) <= TOL


def _check_shock(got, a, dt):
    """One shock step against the oracle.  A pixel may differ only where the
    step's sign decision is a tie: the FP64 second directional derivative
    Inn is within 1e-6 of the image's range of zero (in the near-flat
    regions of a smoothed image its FP32 value is cancellation noise: both
    gradient components sit at the rounding level of the pixel values).
    Such pixels must be rare, and the rest agree to 1e-5."""
    c, flow, Inn = FO.shock_terms(a)
    ref = c - dt * np.copysign(flow, Inn)
    bad = np.abs(got - ref) > 1e-5 * (1 + np.abs(ref))
    if bad.any():
        assert np.abs(Inn[bad]).max() <= 1e-6 * np.abs(a).max(), np.abs(Inn[bad]).max()
    print(f'shock step: {int(bad.sum())} sign ties, max |Inn| there '
          f'{np.abs(Inn[bad]).max() if bad.any() else 0:.2e}', flush=True)
    assert bad.mean() < 1e-3
    assert rel_l2(got[~bad], ref[~bad]) <= TOL
    return int(bad.sum())


def test_prior_2048(cuda):
    """Bilateral (half width 5, sigma_d 3, sigma_r 0.1) then three shock
    steps (dt 0.1) on a 2048^2 latent image, step by step against the
    filters oracle."""
    from vmbd_amd import _lib
    H = W = 2048
    print('case prior 2048^2', flush=True)
    _, x, _ = _scene(53, (H, W), (1, 1), (3, 3), 2)
    xt = _t(x, cuda)
    import torch
    b = torch.empty_like(xt)
    _lib.call('vmbd_bilateral', _lib.ptr(b), _lib.ptr(xt), H, W, 5, 3.0, 0.1, _lib.stream_ptr())
    bn = _n(b)
    assert rel_l2(bn, FO.bilateral(x, 5, 3.0, 0.1)) <= TOL
    a, flips = bn, 0
    for step in range(3):
        o, src = torch.empty_like(xt), _t(a, cuda)
        _lib.call('vmbd_shock_step', _lib.ptr(o), _lib.ptr(src), H, W, 0.1, 1.0, _lib.stream_ptr())
        on = _n(o)
        flips += _check_shock(on, a, 0.1)
        a = on
    print(f'shock sign ties: {flips}', flush=True)


def test_mixed_batch_six_geometries(cuda):
    """Two frames of each of the six mixed-batch geometries in one
    MixedBatch (three streams, one graph replay): every frame's forward and
    exact adjoints against the oracle."""
    import torch
    from vmbd_amd.batch import MixedBatch
    geoms = [((512, 512), (12, 12), (15, 15)), ((512, 512), (6, 6), (41, 41)), ((768, 1024), (8, 8), (25, 25)),
             ((768, 1024), (6, 6), (15, 15)), ((1024, 1024), (10, 10), (41, 41)),
             ((1024, 1024), (12, 12), (25, 25))]
    specs = [geoms[i % 6] for i in range(12)]
    print('case mixed batch six geometries', flush=True)
    mb = MixedBatch(specs, device=cuda, streams=3)
    tiles = {tuple(G['plan'].sx): G['plan'].tile for G in mb.groups}
    assert (256, 512) in {G['plan'].tile for G in mb.groups}, tiles
    rng = np.random.default_rng(54)
    from vmbd_amd import synth
    xs, fs, gs = [], [], []
    for sx, csf, sf in specs:
        xs.append(_f32(synth.piecewise_smooth(sx, rng)))
        fs.append(_f32(synth.motion_psfs(csf[0] * csf[1], sf, rng, steps=sf[0])))
        gs.append(_f32(0.01 * rng.standard_normal((sx[0] - sf[0] + 1, sx[1] - sf[1] + 1))))
    mb.load([_t(v, cuda) for v in xs], [_t(v, cuda) for v in fs], [_t(v, cuda) for v in gs])
    replay = mb.capture()
    for G in mb.groups:
        for k in ('y', 'xb', 'fb'):
            G[k].zero_()
    replay()
    torch.cuda.synchronize()
    for i, (sx, csf, sf) in enumerate(specs):
        owa = O.Win2WinAux(sx, csf, 0.5)
        y, xb, fb = (_n(t) for t in mb.outputs(i))
        e = (rel_l2(y, O.forward(xs[i], fs[i], owa)),
             rel_l2(xb, O.adjoint_x(gs[i], fs[i], owa, sx, exact=True)),
             rel_l2(fb, O.adjoint_f(gs[i], xs[i], owa, sf, exact=True)))
        print(f'frame {i} {sx} {csf} {sf} sw={tuple(owa.sw)} rel L2 {e[0]:.2e} {e[1]:.2e} {e[2]:.2e}', flush=True)
        assert max(e) <= TOL, (i, sx, csf, sf, e)


def _c2_scene():
    """512^2 latent, 8 x 8 patches, 31^2 motion PSFs (15-step walks),
    y = B(x, f) + N(0, 0.01^2) of size 482^2."""
    sx, csf, sf = (512, 512), (8, 8), (31, 31)
    rng, x, f = _scene(55, sx, csf, sf, 15)
    owa = O.Win2WinAux(sx, csf, 0.5)
    y = _f32(O.forward(x, f, owa) + 0.01 * rng.standard_normal((482, 482)))
    return sx, csf, sf, x, f, y, owa


def test_mfbd_frame_512(cuda):
    """One online frame: 20 projected-BB f-iterations (beta 1e-3, warm start
    from flat PSFs) and 20 projected-BB x-iterations with exact adjoints,
    graph captured, against the FP64 oracle frame."""
    from vmbd_amd.mfbd import OnlineMFBD
    sx, csf, sf, x, f, y, owa = _c2_scene()
    print('case mfbd 512^2 8x8 sf31 20+20', flush=True)
    y0 = _f32(O.forward(x, np.full_like(f, 1.0 / (sf[0] * sf[1])), owa))
    m = OnlineMFBD(y0, sf, csf, f_solver='pbb', x_solver='pbb', f_iters=20, x_iters=20, f_beta=1e-3, exact=True,
                   capture=True, device=cuda)
    assert m.plan.tile == (256, 256) and m.plan.sw == (114, 114)
    x0 = _n(m.x)
    x1, f1 = m.step(y)
    x1, f1 = _n(x1), _n(f1)
    fu, xu = _n(m.fu), _n(m.xu)
    f0 = _f32(np.full(f.shape, np.float32(1.0 / (sf[0] * sf[1]))))
    kw = dict(exact=True, x_solver='pbb', xu=xu, x_iters=20, f0=f0)
    xr, fr = M.frame_step(x0, y, owa, sf, fu, 20, 1e-3, **kw)
    old = O.set_fft(WORKERS, True)
    try:
        xs, fs = M.frame_step(x0, y, owa, sf, fu, 20, 1e-3, **kw)
    finally:
        O.set_fft(*old)
    dx, df = rel_l2(xs, xr), rel_l2(fs, fr)
    ex, ef = rel_l2(x1, xr), rel_l2(f1, fr)
    print(f'f: gpu {ef:.2e} fp32-oracle {df:.2e}; x: gpu {ex:.2e} fp32-oracle {dx:.2e}', flush=True)
    assert ef <= max(10 * df, TOL) and ex <= max(10 * dx, TOL)


def test_plain_form_512(cuda):
    """The plain per-operator comparator on the 512^2 online-frame geometry:
    forward and the patch-local transposes against the oracle."""
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.plain import PlainOla
    sx, csf, sf, x, f, y, owa = _c2_scene()
    print('case plain form 512^2', flush=True)
    p = PlainOla(sx, sf, win2winaux(sx, csf, 0.5), device=cuda)
    g = _f32(np.random.default_rng(56).standard_normal(y.shape))
    yy, xb, fb = p.step(_t(x, cuda), _t(f, cuda), _t(g, cuda))
    assert rel_l2(_n(yy), O.forward(x, f, owa)) <= TOL
    assert rel_l2(_n(xb), O.adjoint_x(g, f, owa, sx, exact=False)) <= TOL
    assert rel_l2(_n(fb), O.adjoint_f(g, x, owa, sf, exact=False)) <= TOL
