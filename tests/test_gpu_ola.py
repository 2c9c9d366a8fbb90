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
@pytest.mark.parametrize('image_rows', [0, 1], ids=['rows-by-size', 'rows-shared'])
def test_forward_and_adjoints(geom, mode, image_rows, cuda):
    import torch
    sx, csf, sf, ov = geom
    print(f'case sx={sx} csf={csf} sf={sf} ov={ov} mode={mode} image_rows={image_rows}', flush=True)
    wa, plan, owa = _mk(sx, csf, ov, sf, mode, cuda)
    # 1: the image rows transformed once per patch column (these launches
    # are small enough to take one transform per patch by default)
    plan.set_tuning(image_rows=image_rows)
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


@pytest.mark.parametrize('geom', [((200, 400), (4, 24), (9, 9), 0.75), ((48, 2200), (1, 80), (7, 7), 0.75)],
                         ids=['streamed-rounds', 'wide-rows'])
def test_multi_round_ola(geom, cuda):
    """Output rows covered by more patch pairs than one workgroup round holds
    (60 and 40 pairs for 32 slots): the OLA pass runs several rounds per row,
    gathering each round into LDS column accumulators (rows up to 2048 wide,
    first case) or all rounds at once (the 2194-wide rows of the second).
    Forward and exact adjoints against the oracle."""
    sx, csf, sf, ov = geom
    print(f'case multi-round sx={sx} csf={csf} sf={sf} ov={ov}', flush=True)
    wa, plan, owa = _mk(sx, csf, ov, sf, 'valid', cuda)
    x, f, g = case_inputs(12, sx, csf, sf)
    x32, f32, g32 = (np.float32(a).astype(np.float64) for a in (x, f, g))
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
    assert rel_l2(plan.forward(xt, ft).cpu().numpy(), O.forward(x32, f32, owa)) <= TOL
    xb, fb = plan.adjoint_xf(gt, xt, ft, exact=True)
    assert rel_l2(xb.cpu().numpy(), O.adjoint_x(g32, f32, owa, sx, exact=True)) <= TOL
    assert rel_l2(fb.cpu().numpy(), O.adjoint_f(g32, x32, owa, sf, exact=True)) <= TOL


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


@pytest.mark.parametrize('geom', [GEOMS[1], GEOMS[3], ((300, 700), (2, 3), (25, 25), 0.5)],
                         ids=lambda g: f'{g[0]}-{g[1]}-{g[2]}')
def test_batched_frames_match_single(geom, cuda):
    """B frames in one launch give bitwise the per-frame results, and each
    frame matches the oracle."""
    import torch
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    sx, csf, sf, ov = geom
    print(f'case batched sx={sx} csf={csf} sf={sf}', flush=True)
    B = 3
    wa = win2winaux(sx, csf, ov)
    pb = OlaPlan(sx, sf, wa, device=cuda, max_frames=B)
    p1 = OlaPlan(sx, sf, wa, device=cuda, max_frames=1)
    owa = O.Win2WinAux(sx, csf, ov)
    ins = [case_inputs(20 + b, sx, csf, sf) for b in range(B)]
    X = _t(np.stack([i[0] for i in ins]), cuda)
    F = _t(np.stack([i[1] for i in ins]), cuda)
    G = _t(np.stack([i[2] for i in ins]), cuda)
    Y = pb.forward(X, F)
    XB, FB = pb.adjoint_xf(G, X, F, exact=True)
    XB2 = pb.adjoint_x(G, F, exact=False)
    FB2 = pb.adjoint_f(G, X, exact=False)
    for b in range(B):
        assert torch.equal(Y[b], p1.forward(X[b], F[b]))
        xb1, fb1 = p1.adjoint_xf(G[b], X[b], F[b], exact=True)
        assert torch.equal(XB[b], xb1) and torch.equal(FB[b], fb1)
        assert torch.equal(XB2[b], p1.adjoint_x(G[b], F[b], exact=False))
        assert torch.equal(FB2[b], p1.adjoint_f(G[b], X[b], exact=False))
        x32, f32, g32 = (np.float32(a).astype(np.float64) for a in ins[b])
        assert rel_l2(Y[b].cpu().numpy(), O.forward(x32, f32, owa)) <= TOL
        assert rel_l2(XB[b].cpu().numpy(), O.adjoint_x(g32, f32, owa, sx, exact=True)) <= TOL
        assert rel_l2(FB[b].cpu().numpy(), O.adjoint_f(g32, x32, owa, sf, exact=True)) <= TOL
    # cached spectra, batched
    Fs = pb.spectra('F', F)
    Xs = pb.spectra('X', X)
    # the cached forward transforms the image rows alone, once per patch
    # column; the plain forward (image and PSF rows in one launch at this
    # size) once per patch: the two round differently
    assert rel_l2(pb.forward_cached('F', Fs, X).cpu().numpy(), Y.cpu().numpy()) <= 1e-6
    assert rel_l2(pb.forward_cached('X', Xs, F).cpu().numpy(), Y.cpu().numpy()) <= 1e-6
    assert torch.equal(pb.adjoint_cached('F', Fs, G, exact=True), XB)


def test_row_reuse(cuda):
    """adjoint_xf after forward on the same tensors reuses the row transforms
    (same result bitwise); an in-place change of x disables the reuse."""
    import torch
    sx, csf, sf = (256, 256), (4, 4), (15, 15)
    wa, plan, owa = _mk(sx, csf, 0.5, sf, 'valid', cuda)
    x, f, g = case_inputs(30, sx, csf, sf)
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
    ref_x, ref_f = plan.adjoint_xf(gt, xt, ft, reuse=False)
    plan.forward(xt, ft)
    rx, rf = plan.adjoint_xf(gt, xt, ft)           # reuses
    assert torch.equal(rx, ref_x) and torch.equal(rf, ref_f)
    plan.forward(xt, ft)
    xt.mul_(2.0)                                    # bumps the version counter
    rx, rf = plan.adjoint_xf(gt, xt, ft)            # must recompute
    x2 = np.float32(x).astype(np.float64) * 2.0
    assert rel_l2(rf.cpu().numpy(), O.adjoint_f(np.float32(g).astype(np.float64), x2, owa, sf, exact=True)) <= TOL


@pytest.mark.parametrize('writer', ['cliplower', 'clip', 'edgetaper', 'out'])
def test_row_reuse_after_library_write(writer, cuda):
    """forward(x, f), then a library kernel rewrites x in place through a raw
    pointer, then adjoint_xf(g, x, f): the adjoint must see the new x (the
    wrappers move torch's version counter), checked against the oracle
    evaluated on the modified x."""
    import torch
    from vmbd_amd import gputools
    sx, csf, sf = (256, 256), (4, 4), (15, 15)
    print(f'case reuse after {writer}', flush=True)
    wa, plan, owa = _mk(sx, csf, 0.5, sf, 'valid', cuda)
    x, f, g = case_inputs(32, sx, csf, sf)
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
    plan.forward(xt, ft)
    if writer == 'cliplower':
        gputools.cliplower_GPU(xt, 0.3)
    elif writer == 'clip':
        gputools.clip_GPU(xt, 0.2, 0.7)
    elif writer == 'edgetaper':
        gputools.edgetaper_gpu(xt, (31, 31))
    else:
        # another plan writes x as its output buffer
        wa2, plan2, _ = _mk(sx, csf, 0.5, sf, 'same', cuda)
        plan2.forward(_t(case_inputs(33, sx, csf, sf)[0], cuda), ft, out=xt)
    xmod = xt.cpu().numpy().astype(np.float64)
    assert not np.array_equal(xmod, np.float32(x).astype(np.float64))
    xb, fb = plan.adjoint_xf(gt, xt, ft)
    g32 = np.float32(g).astype(np.float64)
    f32 = np.float32(f).astype(np.float64)
    assert rel_l2(fb.cpu().numpy(), O.adjoint_f(g32, xmod, owa, sf, exact=True)) <= TOL
    assert rel_l2(xb.cpu().numpy(), O.adjoint_x(g32, f32, owa, sx, exact=True)) <= TOL
    torch.cuda.synchronize()


def test_graph_capture_replay(cuda):
    """The forward + both transposes captured in a HIP graph replay to the
    eager results, with eager work between replays."""
    import torch
    sx, csf, sf = (256, 256), (4, 4), (15, 15)
    wa, plan, owa = _mk(sx, csf, 0.5, sf, 'valid', cuda)
    x, f, g = case_inputs(31, sx, csf, sf)
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
    y = torch.empty(plan.sy, device=cuda)
    xb = torch.empty(sx, device=cuda)
    fb = torch.empty((16,) + sf, device=cuda)
    plan.forward(xt, ft, out=y)
    plan.adjoint_xf(gt, xt, ft, out_x=xb, out_f=fb)
    e = (y.clone(), xb.clone(), fb.clone())
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.forward(xt, ft, out=y)
        plan.adjoint_xf(gt, xt, ft, out_x=xb, out_f=fb)
        gr = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gr, stream=s):
            plan.forward(xt, ft, out=y)
            plan.adjoint_xf(gt, xt, ft, out_x=xb, out_f=fb)
    torch.cuda.current_stream().wait_stream(s)
    for it in range(3):
        y.zero_(); xb.zero_(); fb.zero_()
        gr.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, e[0]) and torch.equal(xb, e[1]) and torch.equal(fb, e[2]), it
        _ = (xt * 2).sum().item()  # unrelated eager work between replays


@pytest.mark.parametrize('image_rows', [0, 1], ids=['rows-by-size', 'rows-shared'])
def test_graph_two_streams(image_rows, cuda):
    """Two plans of three frames each, forked over two streams inside one
    graph (the benchmark's step): every replay equals each plan's eager
    results bit for bit, so the plans share no device state."""
    import torch
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    print(f'case two plans on two streams in one graph, image_rows={image_rows}', flush=True)
    sx, csf, sf, B = (256, 256), (4, 4), (15, 15), 3
    wa = win2winaux(sx, csf, 0.5)
    plans = [OlaPlan(sx, sf, wa, mode='valid', device=cuda, max_frames=B) for _ in range(2)]
    for p in plans:
        p.set_tuning(image_rows=image_rows)
    cases = [case_inputs(40 + i, sx, csf, sf) for i in range(2 * B)]
    x, f, g = (_t(np.stack([c[j] for c in cases]), cuda) for j in range(3))
    y = torch.empty((2 * B,) + plans[0].sy, device=cuda)
    xb = torch.empty((2 * B,) + sx, device=cuda)
    fb = torch.empty((2 * B, 16) + sf, device=cuda)
    views = [[t[k * B:(k + 1) * B] for t in (x, f, g, y, xb, fb)] for k in range(2)]

    def batch(k):
        xk, fk, gk, yk, xbk, fbk = views[k]
        plans[k].forward(xk, fk, out=yk)
        plans[k].adjoint_xf(gk, xk, fk, exact=True, out_x=xbk, out_f=fbk)

    for k in range(2):
        batch(k)
    torch.cuda.synchronize()
    e = (y.clone(), xb.clone(), fb.clone())
    s, side = torch.cuda.Stream(), torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for k in range(2):
            batch(k)
        gr = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gr, stream=s):
            cur = torch.cuda.current_stream()
            side.wait_stream(cur)
            batch(0)
            with torch.cuda.stream(side):
                batch(1)
            cur.wait_stream(side)
    torch.cuda.current_stream().wait_stream(s)
    for it in range(3):
        y.zero_(); xb.zero_(); fb.zero_()
        gr.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, e[0]) and torch.equal(xb, e[1]) and torch.equal(fb, e[2]), it
        _ = (x * 2).sum().item()  # unrelated eager work between replays


def test_named_entry_points(cuda):
    """vmbd_windows expands the window classes to win2winaux's ws exactly;
    vmbd_cache_{psf,image}_spectra equal vmbd_spectra; vmbd_pbb_step equals
    vmbd_pbb_stats + vmbd_pbb_update."""
    import torch
    from vmbd_amd import _lib
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    print('case named entry points', flush=True)
    sx, csf, sf = (120, 100), (3, 4), (6, 8)
    wa = win2winaux(sx, csf, 0.5)
    plan = OlaPlan(sx, sf, wa, 'valid')
    assert torch.equal(plan.windows().cpu(), torch.as_tensor(wa.ws))
    x, f, _ = case_inputs(80, sx, csf, sf)
    xt = torch.as_tensor(x.astype(np.float32), device=cuda)
    ft = torch.as_tensor(f.astype(np.float32), device=cuda)
    for kind, name, z in (('F', 'vmbd_cache_psf_spectra', ft), ('X', 'vmbd_cache_image_spectra', xt)):
        ref = plan.spectra(kind, z)
        got = torch.empty_like(ref)
        _lib.call(name, plan._h, 1, _lib.ptr(z), _lib.ptr(got), _lib.stream_ptr())
        assert torch.equal(got, ref), name
    rng = np.random.default_rng(81)
    n = 7777
    a, b, c, d = (torch.as_tensor(rng.standard_normal(n).astype(np.float32), device=cuda) for _ in range(4))
    part = torch.empty(3 * _lib.lib().vmbd_reduce_blocks(n), dtype=torch.float64, device=cuda)
    s1, s2 = torch.empty(1, device=cuda), torch.empty(1, device=cuda)
    o1, o2 = torch.empty_like(a), torch.empty_like(a)
    for it in (1, 2):
        _lib.call('vmbd_pbb_stats', _lib.ptr(a), _lib.ptr(b), _lib.ptr(c), _lib.ptr(d), n, 0, _lib.ptr(part), None,
                  _lib.stream_ptr())
        _lib.call('vmbd_pbb_update', _lib.ptr(o1), _lib.ptr(a), _lib.ptr(c), n, _lib.ptr(part), it, 0, _lib.ptr(s1),
                  _lib.stream_ptr())
        _lib.call('vmbd_pbb_step', _lib.ptr(o2), _lib.ptr(a), _lib.ptr(b), _lib.ptr(c), _lib.ptr(d), n, it, 0,
                  _lib.ptr(part), _lib.ptr(s2), _lib.stream_ptr())
        assert torch.equal(o1, o2) and torch.equal(s1, s2), it


@pytest.mark.parametrize('exact', [False, True])
def test_fused_solver_forms(cuda, exact):
    """forward_cached(sub, mul) = (B u - sub) * mul and adjoint_cached(z, beta)
    = B' g + beta z, for both cache kinds."""
    import torch
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    print(f'case fused solver forms exact={exact}', flush=True)
    sx, csf, sf = (96, 96), (3, 3), (9, 9)
    plan = OlaPlan(sx, sf, win2winaux(sx, csf, 0.5), 'valid')
    x, f, g = case_inputs(90, sx, csf, sf)
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float32), device=cuda)
    xt, ft, gt = t(x), t(f), t(g)
    rng = np.random.default_rng(91)
    sub = t(rng.standard_normal(plan.sy))
    mul = t(rng.random(plan.sy) < 0.8)
    for kind, spec, u, zshape in (('F', plan.spectra('F', ft), xt, sx), ('X', plan.spectra('X', xt), ft, f.shape)):
        y = plan.forward_cached(kind, spec, u)
        assert torch.equal(plan.forward_cached(kind, spec, u, sub=sub, mul=mul), (y - sub) * mul), kind
        assert torch.equal(plan.forward_cached(kind, spec, u, sub=sub), y - sub), kind
        z = t(rng.standard_normal(zshape))
        a = plan.adjoint_cached(kind, spec, gt, exact=exact)
        b = plan.adjoint_cached(kind, spec, gt, exact=exact, z=z, beta=0.37)
        ref = (a.double() + 0.37 * z.double()).cpu().numpy()
        assert rel_l2(b.cpu().numpy(), ref) < 1e-6, kind


@pytest.mark.parametrize('geom', GEOMS, ids=lambda g: f'{g[0]}-{g[1]}-{g[2]}')
def test_separable_windows(geom, cuda):
    """The x-adjoint's summed form (separable windows) against the two-
    dimensional window tables and the oracle, including the canvas's outer
    rows and columns, where every covering window is at the 1e-16 floor and
    the reference weight of each covering patch is 1/count."""
    import copy
    import torch
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    sx, csf, sf, ov = geom
    print(f'case separable sx={sx} csf={csf} sf={sf} ov={ov}', flush=True)
    wa = win2winaux(sx, csf, ov)
    owa = O.Win2WinAux(sx, csf, ov)
    p_sep = OlaPlan(sx, sf, wa, device=cuda)
    p_2d = OlaPlan(sx, sf, wa, device=cuda, separable=False)
    assert p_sep.separable and not p_2d.separable
    x, f, g = case_inputs(7, sx, csf, sf)
    x32, f32, g32 = (np.float32(a).astype(np.float64) for a in (x, f, g))
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)
    for exact in (True, False):
        ref = O.adjoint_x(g32, f32, owa, sx, exact=exact)
        a = p_sep.adjoint_x(gt, ft, exact=exact).cpu().numpy()
        b = p_2d.adjoint_x(gt, ft, exact=exact).cpu().numpy()
        assert rel_l2(a, ref) <= TOL and rel_l2(b, ref) <= TOL, exact
        # outer rows and columns, pixel by pixel
        scale = np.abs(ref).max()
        edge = np.zeros(ref.shape, bool)
        edge[[0, -1], :] = True
        edge[:, [0, -1]] = True
        assert np.abs(a - ref)[edge].max() <= 1e-5 * scale, exact
        xb, fb = p_sep.adjoint_xf(gt, xt, ft, exact=exact)
        assert rel_l2(xb.cpu().numpy(), ref) <= TOL
    # tables that do not factor: the plan keeps its two-dimensional windows
    wb = copy.copy(wa)
    wb.win_table = wa.win_table.copy()
    wb.win_table[0] *= np.float32(1.01)
    p_bad = OlaPlan(sx, sf, wb, device=cuda)
    p_bad2d = OlaPlan(sx, sf, wb, device=cuda, separable=False)
    assert not p_bad.separable
    assert torch.equal(p_bad.adjoint_x(gt, ft), p_bad2d.adjoint_x(gt, ft))


@pytest.mark.parametrize('geom', GEOMS + [((90, 60), (3, 1), (7, 7), 0.5), ((60, 90), (1, 3), (7, 7), 0.5)],
                         ids=lambda g: f'{g[0]}-{g[1]}-{g[2]}')
def test_shared_image_rows(geom, cuda):
    """Separable windows: the image rows are transformed once per patch
    column and scaled per patch row in the column pass, with the floor
    columns' own terms added there.  Forward, f-adjoint and both adjoints
    against the oracle and, pixel by pixel, against the per-patch image rows
    of the two-dimensional path (single patch rows and columns included,
    where one patch column holds both floor columns)."""
    import torch
    from vmbd_amd import _lib
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    sx, csf, sf, ov = geom
    print(f'case shared image rows sx={sx} csf={csf} sf={sf} ov={ov}', flush=True)
    wa = win2winaux(sx, csf, ov)
    owa = O.Win2WinAux(sx, csf, ov)
    p_sep = OlaPlan(sx, sf, wa, device=cuda)
    p_2d = OlaPlan(sx, sf, wa, device=cuda, separable=False)
    assert p_sep.separable and not p_2d.separable
    # small launches keep one transform per patch unless told otherwise
    assert not p_sep.image_rows_shared(1)
    p_sep.set_tuning(image_rows=1)
    assert p_sep.image_rows_shared(1) and p_sep.tuning['image_rows'] == 1
    with pytest.raises(_lib.VmbdError) as e:
        p_2d.set_tuning(image_rows=1)
    assert e.value.code == _lib.VMBD_E_UNSUP
    x, f, g = case_inputs(11, sx, csf, sf)
    x32, f32, g32 = (np.float32(a).astype(np.float64) for a in (x, f, g))
    xt, ft, gt = _t(x, cuda), _t(f, cuda), _t(g, cuda)

    def close(a, b, ref):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert rel_l2(a, ref) <= TOL
        assert np.abs(a - b).max() <= 1e-5 * np.abs(ref).max()

    def run(p):
        y = p.forward(xt, ft)
        fo = p.adjoint_f(gt, xt, exact=True)
        fl = p.adjoint_f(gt, xt, exact=False)
        # the pair reusing the forward's rows
        p.forward(xt, ft)
        xa, fa = p.adjoint_xf(gt, xt, ft, exact=True)
        # cached spectra: the image's (its column spectra from the shared
        # rows) and the PSFs' with the image as the row-pass operand
        Xs, Fs = p.spectra('X', xt), p.spectra('F', ft)
        return (y, fo, fl, xa, fa, p.forward_cached('X', Xs, ft), p.adjoint_cached('X', Xs, gt, exact=True),
                p.adjoint_cached('X', Xs, gt, exact=False), p.forward_cached('F', Fs, xt))

    a, b = run(p_sep), run(p_2d)
    yref = O.forward(x32, f32, owa)
    fref = O.adjoint_f(g32, x32, owa, sf, exact=True)
    flref = O.adjoint_f(g32, x32, owa, sf, exact=False)
    refs = (yref, fref, flref, O.adjoint_x(g32, f32, owa, sx, exact=True), fref, yref, fref, flref, yref)
    for u, v, ref in zip(a, b, refs):
        close(u, v, ref)
    # two wavefronts per column strip: the same bits
    p_sep.set_tuning(col_threads=128, image_rows=1)
    assert all(torch.equal(u, v) for u, v in zip(run(p_sep), a))
    torch.cuda.synchronize()
