"""Mixed-geometry batches (frames of different size, patch grid and PSF
support in one step) and the tile autotuner, against the FP64 oracle and
against single-frame plans."""
import numpy as np
import pytest

from oracle import ola_oracle as O
from tests.helpers import case_inputs, rel_l2

pytestmark = pytest.mark.gpu

SPECS = [((96, 96), (3, 3), (9, 9)), ((70, 90), (2, 3), (7, 5)), ((96, 96), (3, 3), (9, 9)),
         ((120, 100), (3, 4), (6, 8)), ((70, 90), (2, 3), (7, 5))]


def _frames(cuda):
    import torch
    xs, fs, gs = [], [], []
    for i, (sx, csf, sf) in enumerate(SPECS):
        x, f, g = case_inputs(60 + i, sx, csf, sf)
        t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float32), device=cuda)
        xs.append(t(x))
        fs.append(t(f))
        gs.append(t(g))
    return xs, fs, gs


@pytest.mark.parametrize('image_rows', [0, 1], ids=['auto-rows', 'shared-rows'])
@pytest.mark.parametrize('grouped', [True, False], ids=['grouped', 'per-group'])
def test_mixed_batch_vs_oracle_and_single_plans(cuda, grouped, image_rows):
    import torch
    from vmbd_amd.batch import MixedBatch
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    print(f'case mixed batch grouped={grouped} image_rows={image_rows}', flush=True)
    xs, fs, gs = _frames(cuda)
    mb = MixedBatch(SPECS, grouped=grouped)
    assert len(mb.groups) == 3
    for G in mb.groups:  # before the library batches are bound
        G['plan'].set_tuning(image_rows=image_rows)
    # image rows shared per patch column get a row pass of their own in the
    # library's batches (eight launches per step instead of seven); a plan's
    # own step runs them beside the PSF rows in one launch
    shared = [G['plan'].image_rows_shared(len(G['idx'])) for G in mb.groups]
    tiles = {G['plan'].tile for G in mb.groups}
    if grouped:
        # grouped over three streams: one batch per tile, 7 launches each
        # plus one where a group of the tile shares its image rows
        assert mb.launches == sum(7 + any(sh for G, sh in zip(mb.groups, shared) if G['plan'].tile == t)
                                  for t in tiles)
    else:
        assert mb.launches == 7 * len(mb.groups)
    if grouped:
        # on one stream one batch of all groups: one launch per pass and
        # distinct tile (width for the five row, overlap-add and stack
        # passes and the shared image rows, height for the two column passes)
        m1 = MixedBatch(SPECS, grouped=True, streams=1)
        for G in m1.groups:
            G['plan'].set_tuning(image_rows=image_rows)
        n_tw = len({G['plan'].tile[1] for G in m1.groups})
        n_th = len({G['plan'].tile[0] for G in m1.groups})
        n_x = len({G['plan'].tile[1] for G in m1.groups if G['plan'].image_rows_shared(len(G['idx']))})
        assert m1.launches == 5 * n_tw + n_x + 2 * n_th
    mb.load(xs, fs, gs)
    mb.step()
    torch.cuda.synchronize()
    for i, (sx, csf, sf) in enumerate(SPECS):
        y, xb, fb = mb.outputs(i)
        owa = O.Win2WinAux(sx, csf, 0.5)
        x64, f64, g64 = (t.cpu().numpy().astype(np.float64) for t in (xs[i], fs[i], gs[i]))
        assert rel_l2(y.cpu().numpy(), O.forward(x64, f64, owa)) < 1e-5, i
        assert rel_l2(xb.cpu().numpy(), O.adjoint_x(g64, f64, owa, sx, exact=True)) < 1e-5, i
        assert rel_l2(fb.cpu().numpy(), O.adjoint_f(g64, x64, owa, sf, exact=True)) < 1e-5, i
        # batching changes nothing: the single-frame plan gives the same bits
        p1 = OlaPlan(sx, sf, win2winaux(sx, csf, 0.5), 'valid')
        p1.set_tuning(image_rows=image_rows)
        assert torch.equal(p1.forward(xs[i], fs[i]), y), i
        xb1, fb1 = p1.adjoint_xf(gs[i], xs[i], fs[i], exact=True)
        assert torch.equal(xb1, xb) and torch.equal(fb1, fb), i


@pytest.mark.parametrize('streams,grouped', [(1, False), (2, False), (3, False), (1, True), (3, True)])
def test_mixed_batch_graph_replay(cuda, streams, grouped):
    """The captured step equals the eager one, also with the groups forked
    over several streams inside the graph, and with the grouped launches."""
    import torch
    from vmbd_amd.batch import MixedBatch
    print(f'case mixed batch graph, {streams} streams, grouped={grouped}', flush=True)
    xs, fs, gs = _frames(cuda)
    mb = MixedBatch(SPECS, streams=streams, grouped=grouped)
    mb.load(xs, fs, gs)
    mb.step()
    torch.cuda.synchronize()
    ref = [tuple(t.clone() for t in mb.outputs(i)) for i in range(len(SPECS))]
    replay = mb.capture()
    for rep in range(3):
        for G in mb.groups:
            G['y'].zero_()
            G['xb'].zero_()
            G['fb'].zero_()
        torch.cuda.synchronize()
        replay()
        torch.cuda.synchronize()
        for i in range(len(SPECS)):
            for a, b in zip(mb.outputs(i), ref[i]):
                assert torch.equal(a, b), (rep, i)


def test_autotuned_tile_gives_same_result(cuda, tmp_path, monkeypatch):
    import torch
    from vmbd_amd import autotune
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    print('case autotune', flush=True)
    monkeypatch.setattr(autotune, 'CACHE', str(tmp_path / 'tiles.json'))
    sx, csf, sf = (96, 96), (3, 3), (9, 9)
    wa = win2winaux(sx, csf, 0.5)
    tile, times = autotune.best_tile(sx, sf, wa, frames=2)
    assert tile in autotune.candidates(np.broadcast_to(wa.sw, (2,)), sf) and len(times) == 4
    assert tile == autotune.best_tile(sx, sf, wa, frames=2)[0]          # cached
    x, f, g = case_inputs(70, sx, csf, sf)
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float32), device=cuda)
    owa = O.Win2WinAux(sx, csf, 0.5)
    x64, f64 = (np.asarray(a, np.float32).astype(np.float64) for a in (x, f))
    for tl in autotune.candidates(np.broadcast_to(wa.sw, (2,)), sf):
        p = OlaPlan(sx, sf, wa, 'valid', tile=tl)
        assert rel_l2(p.forward(t(x), t(f)).cpu().numpy(), O.forward(x64, f64, owa)) < 1e-5, tl


def test_launch_shapes_same_bits_and_tuner(cuda, tmp_path, monkeypatch):
    """Every launch shape a plan accepts (overlap-add and column-pass
    workgroup sizes) gives the same bits; shapes a geometry cannot use are
    refused; the tuner's search covers tile x shape and its choice is one of
    the accepted configurations."""
    import torch
    from vmbd_amd import _lib, autotune
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.ola import OlaPlan
    print('case launch shapes', flush=True)
    monkeypatch.setattr(autotune, 'CACHE', str(tmp_path / 'tiles.json'))
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float32), device=cuda)
    for sx, csf, sf in (((256, 256), (4, 4), (15, 15)), ((200, 400), (4, 24), (9, 9))):
        wa = win2winaux(sx, csf, 0.5)
        x, f, g = case_inputs(71, sx, csf, sf)
        ref = None
        accepted = 0
        for shape in autotune.SHAPES:
            p = OlaPlan(sx, sf, wa, 'valid')
            try:
                p.set_tuning(*shape)
            except _lib.VmbdError as e:
                assert e.code == _lib.VMBD_E_UNSUP and shape[0] in (128, 192), (shape, e)
                continue
            assert p.tuning == dict(ola_threads=shape[0], col_threads=shape[1], image_rows=0)
            out = (p.forward(t(x), t(f)),) + p.adjoint_xf(t(g), t(x), t(f))
            if ref is None:
                ref = out
            assert all(torch.equal(a, b) for a, b in zip(out, ref)), (sx, shape)
            accepted += 1
        assert accepted >= 4, sx
    sx, csf, sf = (256, 256), (4, 4), (15, 15)
    wa = win2winaux(sx, csf, 0.5)
    cfg, times = autotune.best_config(sx, sf, wa, frames=2)
    assert len(times) >= 8
    assert autotune.best_config(sx, sf, wa, frames=2)[0] == cfg          # cached
    p = OlaPlan(sx, sf, wa, 'valid', tile=cfg['tile'])
    p.set_tuning(cfg['ola_threads'], cfg['col_threads'])
