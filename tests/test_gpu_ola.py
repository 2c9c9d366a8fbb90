"""GPU parity of the EFF/OLA operator (HIP kernels through the C ABI) against
the FP64 oracle.  Tolerance: relative L2 <= 1e-5 for every output."""
import numpy as np
import pytest

from oracle import ola_oracle as O
from tests.helpers import GEOMS, case_inputs, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _mk(sx, csf, overlap, sf, mode, cuda):
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    wa = win2winaux(sx, csf, overlap)
    plan = OlaPlan(sx, sf, wa, mode=mode, device=cuda)
    owa = O.Win2WinAux(sx, csf, overlap)
    return wa, plan, owa


def _t(a, cuda):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=cuda)


@pytest.mark.parametrize('geom', GEOMS, ids=lambda g: f'{g[0]}-{g[1]}-{g[2]}')
@pytest.mark.parametrize('mode', ['valid', 'same', 'full'])
def test_forward_and_adjoints(geom, mode, cuda):
    import torch
    sx, csf, sf, ov = geom
    print(f'case sx={sx} csf={csf} sf={sf} ov={ov} mode={mode}', flush=True)
    wa, plan, owa = _mk(sx, csf, ov, sf, mode, cuda)
    x, f, g = case_inputs(1, sx, csf, sf, mode)
    # the reference computes in float32: feed the oracle the float32-rounded inputs
    x32, f32, g32 = (np.float32(a).astype(np.float64) for a in (x, f, g))
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)

    y = plan.forward(xt, ft)
    y_ref = O.forward(x32, f32, owa, mode)
    assert rel_l2(y.cpu().numpy(), y_ref) <= TOL

    if mode == 'full':
        # the reference's transposes exist for valid/same only: check the
        # full-mode transposes against the forward by the dot-product identity
        xb, fb = plan.adjoint_xf(gt, xt, ft, exact=True)
        lhs = float(np.dot(y.cpu().numpy().ravel().astype(np.float64), g32.ravel()))
        rx = float(np.dot(x32.ravel(), xb.cpu().numpy().ravel().astype(np.float64)))
        rf = float(np.dot(f32.ravel(), fb.cpu().numpy().ravel().astype(np.float64)))
        assert abs(lhs - rx) <= 1e-5 * (abs(lhs) + np.linalg.norm(y_ref) * np.linalg.norm(g32))
        assert abs(lhs - rf) <= 1e-5 * (abs(lhs) + np.linalg.norm(y_ref) * np.linalg.norm(g32))
        torch.cuda.synchronize()
        return
    for exact in (True, False):
        xb = plan.adjoint_x(gt, ft, exact=exact)
        xb_ref = O.adjoint_x(g32, f32, owa, sx, mode, exact=exact)
        assert rel_l2(xb.cpu().numpy(), xb_ref) <= TOL, ('adjoint_x', exact)
        fb = plan.adjoint_f(gt, xt, exact=exact)
        fb_ref = O.adjoint_f(g32, x32, owa, sf, mode, exact=exact)
        assert rel_l2(fb.cpu().numpy(), fb_ref) <= TOL, ('adjoint_f', exact)
        xb2, fb2 = plan.adjoint_xf(gt, xt, ft, exact=exact)
        assert rel_l2(xb2.cpu().numpy(), xb_ref) <= TOL
        assert rel_l2(fb2.cpu().numpy(), fb_ref) <= TOL
    torch.cuda.synchronize()


@pytest.mark.parametrize('geom', GEOMS[:4], ids=lambda g: f'{g[0]}-{g[1]}-{g[2]}')
def test_cached_spectra(geom, cuda):
    """OlaGPU semantics: F instance caches F^, X instance caches X^."""
    sx, csf, sf, ov = geom
    print(f'case sx={sx} csf={csf} sf={sf}', flush=True)
    wa, plan, owa = _mk(sx, csf, ov, sf, 'valid', cuda)
    x, f, g = case_inputs(2, sx, csf, sf)
    x32, f32, g32 = (np.float32(a).astype(np.float64) for a in (x, f, g))
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
    Fs = plan.spectra('F', ft)
    Xs = plan.spectra('X', xt)
    y_ref = O.forward(x32, f32, owa)
    assert rel_l2(plan.forward_cached('F', Fs, xt).cpu().numpy(), y_ref) <= TOL
    assert rel_l2(plan.forward_cached('X', Xs, ft).cpu().numpy(), y_ref) <= TOL
    for exact in (False, True):
        xb_ref = O.adjoint_x(g32, f32, owa, sx, exact=exact)
        fb_ref = O.adjoint_f(g32, x32, owa, sf, exact=exact)
        assert rel_l2(plan.adjoint_cached('F', Fs, gt, exact=exact).cpu().numpy(), xb_ref) <= TOL
        assert rel_l2(plan.adjoint_cached('X', Xs, gt, exact=exact).cpu().numpy(), fb_ref) <= TOL


def test_c3_geometry(cuda):
    """1024^2 image, 10x10 patches, 41x41 PSFs (T = 256)."""
    sx, csf, sf = (1024, 1024), (10, 10), (41, 41)
    print('case C3 1024^2 10x10 sf41', flush=True)
    wa, plan, owa = _mk(sx, csf, 0.5, sf, 'valid', cuda)
    assert plan.tile == (256, 256)
    x, f, g = case_inputs(3, sx, csf, sf)
    x32, f32, g32 = (np.float32(a).astype(np.float64) for a in (x, f, g))
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
    y = plan.forward(xt, ft)
    assert rel_l2(y.cpu().numpy(), O.forward(x32, f32, owa)) <= TOL
    xb, fb = plan.adjoint_xf(gt, xt, ft, exact=True)
    assert rel_l2(xb.cpu().numpy(), O.adjoint_x(g32, f32, owa, sx, exact=True)) <= TOL
    assert rel_l2(fb.cpu().numpy(), O.adjoint_f(g32, x32, owa, sf, exact=True)) <= TOL


def test_tile_512(cuda):
    """T = 512 tiles (sw + sf - 1 > 256) and rectangular tiles."""
    for sx, csf, sf in (((600, 600), (3, 3), (41, 41)), ((300, 700), (2, 3), (25, 25))):
        print(f'case sx={sx} csf={csf} sf={sf}', flush=True)
        wa, plan, owa = _mk(sx, csf, 0.5, sf, 'valid', cuda)
        x, f, g = case_inputs(4, sx, csf, sf)
        x32, f32, g32 = (np.float32(a).astype(np.float64) for a in (x, f, g))
        xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
        assert rel_l2(plan.forward(xt, ft).cpu().numpy(), O.forward(x32, f32, owa)) <= TOL
        xb, fb = plan.adjoint_xf(gt, xt, ft, exact=True)
        assert rel_l2(xb.cpu().numpy(), O.adjoint_x(g32, f32, owa, sx, exact=True)) <= TOL
        assert rel_l2(fb.cpu().numpy(), O.adjoint_f(g32, x32, owa, sf, exact=True)) <= TOL


def test_deterministic(cuda):
    """Two runs give bitwise identical outputs (no atomics in the OLA)."""
    sx, csf, sf = (256, 256), (4, 4), (15, 15)
    wa, plan, owa = _mk(sx, csf, 0.5, sf, 'valid', cuda)
    x, f, g = case_inputs(5, sx, csf, sf)
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
    y1 = plan.forward(xt, ft).cpu().numpy()
    y2 = plan.forward(xt, ft).cpu().numpy()
    assert np.array_equal(y1, y2)
    a1 = plan.adjoint_xf(gt, xt, ft)
    a2 = plan.adjoint_xf(gt, xt, ft)
    assert np.array_equal(a1[0].cpu().numpy(), a2[0].cpu().numpy())
    assert np.array_equal(a1[1].cpu().numpy(), a2[1].cpu().numpy())
