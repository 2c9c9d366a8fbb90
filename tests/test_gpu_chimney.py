"""End to end on the reference's own frame sequence (tests/golden/chimney):
the online loop of vmbd_chimney.py over all 100 frames.

The quality figure is the PSNR against groundtruth.png after the best
affine intensity fit and integer shift (|shift| <= 3) over the interior,
since the loop's output is normalised by its clip level and may be offset
by the PSF centring.  On this figure the initial estimate (mean of frames
1-19 / 20) scores 33.17 dB and the reference's own result.png 33.81 dB.

The reference's loop, restated in FP64 with its own parameters
(tests/golden/make_chimney_oracle_psnr.py), peaks near frame 5 and then
loses quality: 35.0 dB at frame 10, 29.9 at frame 100 (29.8 with the
reference's x-transpose reading past each patch).  Its 100-frame result is
below result.png, so result.png is not what the shipped loop produces, and
the bars are against the FP64 trajectory of the same algorithm
(tests/golden/chimney_oracle_psnr.json): the L-BFGS-B loop must stay within
0.5 dB of it, or above, at every checkpoint, and beat result.png after 10
frames; the device-PBB loop (20 iterations per frame) must beat the initial
estimate after 10 frames and end within 0.5 dB of the FP64 loop."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'chimney')


def aligned_psnr(a, b):
    h, w = min(a.shape[0], b.shape[0]), min(a.shape[1], b.shape[1])
    a = np.asarray(a, dtype=np.float64)[:h, :w]
    b = np.asarray(b, dtype=np.float64)[:h, :w]
    best = -np.inf
    for dy in range(-3, 4):
        for dx in range(-3, 4):
            aa = np.roll(np.roll(a, dy, 0), dx, 1)[5:-5, 5:-5]
            bb = b[5:-5, 5:-5]
            A = np.stack([aa.ravel(), np.ones(aa.size)], 1)
            c = np.linalg.lstsq(A, bb.ravel(), rcond=None)[0]
            best = max(best, 10 * np.log10(1.0 / np.mean(((A @ c).reshape(aa.shape) - bb) ** 2)))
    return best


@pytest.fixture(scope='module')
def reference_images():
    from vmbd_amd.imagetools import imread
    from vmbd_amd.mfbd import initial_latent
    gt = imread(os.path.join(DATA, 'groundtruth.png'))
    rs = imread(os.path.join(DATA, 'result.png'))
    y0 = initial_latent([imread(os.path.join(DATA, 'y_%04d.png' % i)) for i in range(1, 20)], 20)
    return gt, aligned_psnr(rs, gt), aligned_psnr(y0, gt)


CHECK = (10, 30, 50, 75, 100)


@pytest.fixture(scope='module')
def oracle_psnr():
    import json
    with open(os.path.join(os.path.dirname(DATA), 'chimney_oracle_psnr.json')) as fh:
        return {int(k): v for k, v in json.load(fh)['fixed'].items()}


def test_chimney_lbfgsb(cuda, reference_images, oracle_psnr):
    from vmbd_amd.chimney import run
    gt, p_result, p_init = reference_images
    print('case chimney lbfgsb 100 frames', flush=True)
    xi, fs, sec = run(DATA, N=100, f_solver='lbfgsb', optiter=500, snapshots=CHECK)
    np.testing.assert_allclose(fs.sum(axis=(1, 2)), 1.0, rtol=1e-5)
    p = {k: aligned_psnr(v, gt) for k, v in run.snapshots.items()}
    print('psnr ' + ' '.join(f'{k}:{p[k]:.3f}/{oracle_psnr[k]:.3f}' for k in CHECK) +
          f' (result.png {p_result:.3f}, initial {p_init:.3f}), {sec:.1f} s', flush=True)
    assert p[10] > p_result
    assert p[10] > p_init + 1.0
    for k in CHECK:
        assert p[k] >= oracle_psnr[k] - 0.5, k


def test_chimney_device_pbb(cuda, reference_images, oracle_psnr):
    from vmbd_amd.chimney import run
    gt, p_result, p_init = reference_images
    print('case chimney device pbb 100 frames', flush=True)
    xi, fs, sec = run(DATA, N=100, f_solver='pbb', f_iters=20, capture=True, snapshots=CHECK)
    p = {k: aligned_psnr(v, gt) for k, v in run.snapshots.items()}
    print('psnr ' + ' '.join(f'{k}:{p[k]:.3f}/{oracle_psnr[k]:.3f}' for k in CHECK) + f' (initial {p_init:.3f})',
          flush=True)
    assert p[10] > p_init
    assert p[100] >= oracle_psnr[100] - 0.5
