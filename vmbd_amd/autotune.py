"""Tile autotuner for the fused operator.

Any FFT tile at least sw+sf-1 per dimension gives the same result up to
rounding (olaGPU.py:139-141 picks the smallest power of two), so the tile
is free to choose.  For a geometry, ``best_tile`` times one forward plus
adjoint pair of B frames for each candidate (the smallest power of two and
the next one per dimension, within the kernels' 16..512 range) and keeps the
fastest; results are cached per geometry in a JSON file.
"""
import json
import os

import numpy as np
import torch

from .imagetools import fft_tile

CACHE = os.path.join(os.path.expanduser('~'), '.cache', 'vmbd_amd', 'tiles.json')


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


def make_plan(sx, sf, winaux, tile, frames=1, device=None):
    """The plan of a candidate tile, or None when the library refuses the
    tile for this geometry (VMBD_E_UNSUP / VMBD_E_ARG at plan creation, e.g.
    an overlap-add that needs more LDS than a workgroup has); any other
    failure propagates."""
    from . import _lib
    from .ola import OlaPlan
    dev = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
    try:
        return OlaPlan(sx, sf, winaux, 'valid', tile=tile, device=dev, max_frames=frames)
    except _lib.VmbdError as e:
        if e.code in (_lib.VMBD_E_UNSUP, _lib.VMBD_E_ARG):
            return None
        raise


def time_tile(sx, sf, winaux, tile, frames=1, reps=5, device=None, plan=None):
    """Milliseconds of forward + adjoint_xf of `frames` frames on `tile`."""
    dev = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
    if plan is None:
        plan = make_plan(sx, sf, winaux, tile, frames, dev)
        if plan is None:
            raise ValueError(f'tile {tile} is not supported for this geometry')
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.rand((frames,) + tuple(sx), generator=g, device=dev)
    f = torch.rand((frames, plan.P) + tuple(sf), generator=g, device=dev)
    r = torch.rand((frames,) + plan.sy, generator=g, device=dev)
    y = torch.empty_like(r)
    xb = torch.empty_like(x)
    fb = torch.empty_like(f)

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


def best_tile(sx, sf, winaux, frames=1, overlap=None, use_cache=True, device=None):
    """Fastest tile for the geometry (cached); returns (tile, {tile: ms})."""
    csf = winaux.csf
    ov = overlap if overlap is not None else getattr(winaux, 'overlap', 0.5)
    dev = torch.device(device) if device is not None else torch.device('cuda', torch.cuda.current_device())
    key = _key(sx, sf, csf, ov, frames, dev)
    cache = _load() if use_cache else {}
    if key in cache:
        return tuple(cache[key]['tile']), {tuple(map(int, k.split('x'))): v for k, v in cache[key]['ms'].items()}
    sw = np.broadcast_to(np.asarray(winaux.sw), (2,))
    times = {}
    for t in candidates(sw, sf):
        plan = make_plan(sx, sf, winaux, t, frames, dev)
        if plan is None:
            continue            # a tile the plan rejects (LDS limits) is not a candidate
        times[t] = time_tile(sx, sf, winaux, t, frames, device=dev, plan=plan)
        del plan
    if not times:
        raise RuntimeError('no tile candidate is supported for this geometry')
    best = min(times, key=times.get)
    if use_cache:
        cache[key] = {'tile': list(best), 'ms': {f'{a}x{b}': v for (a, b), v in times.items()}}
        _store(cache)
    return best, times
