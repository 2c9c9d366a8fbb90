"""Online multi-frame blind deconvolution of a frame sequence on disk: the
application script of the reference (vmbd_chimney.py:1-185) on the drop-in.

    python -m vmbd_amd.chimney --data DIR [--frames 100] [--out DIR]

Frames are DIR/y_0001.png ... (8-bit); the latent image starts as the mean
of the first N0-1 frames divided by N0, replicate-padded by sf-1
(vmbd_chimney.py:89-100); each frame runs one PSF estimate and one
multiplicative latent-image step (vmbd_chimney.py:110-147).  With --out the
observed frame, the latent image (cropped by floor(sf/2) and divided by its
clip level) and the PSF mosaic are written per frame as the reference does
(vmbd_chimney.py:153-163).
"""
import argparse
import os
import time

import numpy as np

from .imagetools import gridF, imread, imwrite
from .mfbd import OnlineMFBD, initial_latent


def frame_path(datapath, i, pattern='y_%04d.png'):
    return os.path.join(datapath, pattern % i)


def run(datapath, N=100, N0=20, sf=(20, 20), csf=(3, 3), overlap=0.5, f_alpha=0., f_beta=0.1, optiter=500,
        tol=1e-10, f_solver='lbfgsb', f_iters=20, f_init='warm', capture=False, respath=None, verbose=False,
        snapshots=()):
    """Process frames 1..N; returns (latent image cropped by floor(sf/2) and
    divided by its clip level, PSFs, seconds).  ``snapshots``: frame numbers
    whose latent image (same crop and scale) is kept in ``run.snapshots``."""
    sf = np.broadcast_to(np.asarray(sf, dtype=int), (2,))
    load = lambda i: imread(frame_path(datapath, i))
    y0 = initial_latent([load(i) for i in range(1, N0)], N0)
    m = OnlineMFBD(y0, sf, csf, overlap, f_solver=f_solver, x_solver='rl', f_alpha=f_alpha, f_beta=f_beta,
                   optiter=optiter, tol=tol, f_iters=f_iters, capture=capture, f_init=f_init)
    sf2 = sf // 2
    if respath:
        os.makedirs(respath, exist_ok=True)
    t0 = time.perf_counter()
    xi = fs = None
    run.snapshots = {}
    for i in range(1, N + 1):
        y = load(i)
        x, f = m.step(y)
        if i in snapshots:
            xh = x.cpu().numpy()
            run.snapshots[i] = xh[sf2[0]:-sf2[0], sf2[1]:-sf2[1]] / float(m.xmax.item())
        if respath or i == N:
            x_max = float(m.xmax.item())
            xh = x.cpu().numpy()
            xi = xh[sf2[0]:-sf2[0], sf2[1]:-sf2[1]] / x_max
            fs = f.cpu().numpy()
        if respath:
            imwrite(y, os.path.join(respath, 'y_%04d.png' % i))
            imwrite(xi, os.path.join(respath, 'x_%04d.png' % i))
            fg = gridF(fs, csf)
            imwrite(fg / fg.max(), os.path.join(respath, 'f_%04d.png' % i))
        if verbose:
            print('Processing frame %d/%d' % (i, N), flush=True)
    return xi, fs, time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--data', required=True)
    ap.add_argument('--frames', type=int, default=100)
    ap.add_argument('--n0', type=int, default=20)
    ap.add_argument('--sf', type=int, nargs=2, default=(20, 20))
    ap.add_argument('--csf', type=int, nargs=2, default=(3, 3))
    ap.add_argument('--overlap', type=float, default=0.5)
    ap.add_argument('--f-alpha', type=float, default=0.)
    ap.add_argument('--f-beta', type=float, default=0.1)
    ap.add_argument('--optiter', type=int, default=500)
    ap.add_argument('--tol', type=float, default=1e-10)
    ap.add_argument('--f-solver', choices=('lbfgsb', 'pbb'), default='lbfgsb')
    ap.add_argument('--f-iters', type=int, default=20)
    ap.add_argument('--f-init', choices=('warm', 'xty'), default='warm')
    ap.add_argument('--capture', action='store_true')
    ap.add_argument('--out', default=None)
    a = ap.parse_args()
    xi, fs, dt = run(a.data, a.frames, a.n0, a.sf, a.csf, a.overlap, a.f_alpha, a.f_beta, a.optiter, a.tol,
                     a.f_solver, a.f_iters, a.f_init, a.capture, a.out, verbose=True)
    print('%d frames in %.2f s (%.2f frames/s)' % (a.frames, dt, a.frames / dt))


if __name__ == '__main__':
    main()
