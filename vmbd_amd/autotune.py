"""Autotuner for the fused operator's launch configuration.

Three dimensions, none of which changes a result (every configuration gives
the same operator up to rounding of the FFT size, and the same bits for a
given tile):

* the FFT tile: any tile at least sw+sf-1 per dimension gives the same
  result up to rounding (olaGPU.py:139-141 picks the smallest power of two);
  candidates are the smallest power of two and the next one per dimension,
  within the kernels' 16..512 range;
* the overlap-add passes' workgroup size (ola_threads: 128 / 192 / 256, the
  first two when every output row fits one round), i.e. how many output
  rows share a compute unit;
* the column passes' workgroup size (col_threads: 128 / 256), i.e. how many
  wavefronts carry one strip of a patch's columns (two or four at a
  256-point tile) and how much LDS a workgroup holds.

``best_config`` times a workload of B frames for each candidate (the
batched operator, or the cached-operand gradients of the online loop's
solvers, whose best shapes differ) and keeps the fastest; results are
cached per geometry, frame count, workload, GPU architecture and library
version in a JSON file.
"""
import json
import os

import numpy as np
import torch

from .imagetools import fft_tile

CACHE = os.path.join(os.path.expanduser('~'), '.cache', 'vmbd_amd', 'tiles.json')

# (ola_threads, col_threads) tried per tile; 0 = the library's own choice
SHAPES = [(0, 0), (256, 0), (128, 0), (192, 0), (0, 128), (256, 128)]


def _key(sx, sf, csf, overlap, frames, device):
    """Cache key: the geometry, the frame count, the device's architecture
    and the library version (a timing taken on another GPU or with other
    kernels is not reused)."""
    from . import _lib
    arch = torch.cuda.get_device_properties(device).gcnArchName
    ver = _lib.lib().vmbd_version()
    return f'{sx[0]}x{sx[1]}/{sf[0]}x{sf[1]}/{csf[0]}x{csf[1]}/{overlap}/{frames}/{arch}/v{ver}'


def candidates(sw, sf):
    t0 = fft_tile(sw, sf)
    out = []
    for a in (t0[0], 2 * t0[0]):
        for b in (t0[1], 2 * t0[1]):
            if a <= 512 and b <= 512:
                out.append((a, b))
    return out


def _load():
    try:
        with open(CACHE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(d):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE + '.tmp', 'w') as fh:
            json.dump(d, fh)
        os.replace(CACHE + '.tmp', CACHE)
    except OSError:
        pass


def _refused(e):
    """A configuration the library refuses for this geometry (as opposed to
    a failure): VMBD_E_UNSUP / VMBD_E_ARG at plan creation or tuning."""
    from . import _lib
    return isinstance(e, _lib.VmbdError) and e.code in (_lib.VMBD_E_UNSUP, _lib.VMBD_E_ARG)


def make_plan(sx, sf, winaux, tile, frames=1, device=None, shape=(0, 0)):
    """The plan of a candidate (tile, launch shape), or None when the library
    refuses it for this geometry (e.g. an overlap-add that needs more LDS
    than a workgroup has, or a row that does not fit one round of
    ola_threads); any other failure propagates."""
    from .ola import OlaPlan
    dev = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
    try:
        plan = OlaPlan(sx, sf, winaux, 'valid', tile=tile, device=dev, max_frames=frames)
        plan.set_tuning(*shape)
        return plan
    except Exception as e:  # noqa: BLE001 - re-raised unless it is a refusal
        if _refused(e):
            return None
        raise


WORKLOADS = ('step', 'cached')


def time_plan(plan, frames=1, reps=5, workload='step'):
    """Milliseconds of `frames` frames of a workload on the plan: 'step' =
    forward + adjoint_xf with both operands fresh (the batched operator);
    'cached' = one gradient of each unknown with the other cached, forward
    and adjoint (the iterations of the online loop's solvers)."""
    if workload not in WORKLOADS:
        raise ValueError(f'workload must be one of {WORKLOADS}')
    dev = plan.device
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.rand((frames,) + tuple(plan.sx), generator=g, device=dev)
    f = torch.rand((frames, plan.P) + tuple(plan.sf), generator=g, device=dev)
    r = torch.rand((frames,) + plan.sy, generator=g, device=dev)
    y = torch.empty_like(r)
    xb = torch.empty_like(x)
    fb = torch.empty_like(f)
    if workload == 'cached':
        sF, sX = plan.spectra('F', f), plan.spectra('X', x)

        def step():
            plan.forward_cached('X', sX, f, out=y)
            plan.adjoint_cached('X', sX, r, exact=True, out=fb)
            plan.forward_cached('F', sF, x, out=y)
            plan.adjoint_cached('F', sF, r, exact=True, out=xb)
    else:
        def step():
            plan.forward(x, f, out=y)
            plan.adjoint_xf(r, x, f, out_x=xb, out_f=fb)

    step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        step()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def time_tile(sx, sf, winaux, tile, frames=1, reps=5, device=None, shape=(0, 0)):
    """Milliseconds of forward + adjoint_xf of `frames` frames on `tile` with
    the launch shape (ola_threads, col_threads)."""
    plan = make_plan(sx, sf, winaux, tile, frames, device, shape)
    if plan is None:
        raise ValueError(f'tile {tile} / shape {shape} is not supported for this geometry')
    return time_plan(plan, frames, reps)


def _cfg_key(tile, shape):
    return f'{tile[0]}x{tile[1]}/{shape[0]}/{shape[1]}'


def best_config(sx, sf, winaux, frames=1, overlap=None, use_cache=True, device=None, shapes=None, workload='step'):
    """Fastest (tile, ola_threads, col_threads) for the geometry and workload
    (time_plan; cached); returns (dict(tile, ola_threads, col_threads),
    {config key: ms})."""
    csf = winaux.csf
    ov = overlap if overlap is not None else getattr(winaux, 'overlap', 0.5)
    dev = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
    shapes = list(shapes or SHAPES)
    key = _key(sx, sf, csf, ov, frames, dev) + f'/{workload}/' + ';'.join(f'{a},{b}' for a, b in shapes)
    cache = _load() if use_cache else {}
    if key in cache and 'best' in cache[key]:
        b = cache[key]['best']
        return dict(tile=tuple(b['tile']), ola_threads=b['ola_threads'], col_threads=b['col_threads']), cache[key]['ms']
    sw = np.broadcast_to(np.asarray(winaux.sw), (2,))
    times, cfgs = {}, {}
    for t in candidates(sw, sf):
        for shape in shapes:
            plan = make_plan(sx, sf, winaux, t, frames, dev, shape)
            if plan is None:
                continue        # refused for this geometry: not a candidate
            k = _cfg_key(t, shape)
            times[k] = time_plan(plan, frames, workload=workload)
            cfgs[k] = dict(tile=tuple(t), ola_threads=shape[0], col_threads=shape[1])
            del plan
    if not times:
        raise RuntimeError('no tile candidate is supported for this geometry')
    best = cfgs[min(times, key=times.get)]
    if use_cache:
        cache[key] = {'best': dict(best, tile=list(best['tile'])), 'ms': times}
        _store(cache)
    return best, times


def best_tile(sx, sf, winaux, frames=1, overlap=None, use_cache=True, device=None):
    """Fastest tile with the library's own launch shapes (cached); returns
    (tile, {tile: ms})."""
    cfg, times = best_config(sx, sf, winaux, frames, overlap, use_cache, device, shapes=[(0, 0)])
    out = {}
    for k, v in times.items():
        a, b = (int(q) for q in k.split('/')[0].split('x'))
        out[(a, b)] = v
    return cfg['tile'], out
