"""FP64 trajectory of the reference's chimney loop on the CPU, written to
tests/golden/chimney_oracle_psnr.json (the bars of tests/test_gpu_chimney.py).

The loop of vmbd_chimney.py:33-147 with its parameters (N0 = 20, sf 20x20,
3x3 patches, overlap 0.5, f-step by L-BFGS-B on the objective and gradient
of olaGPU.py:756-811 with alpha = 0, beta = 0.1, 500 evaluations, patch-local
transposes, PSF normalisation, multiplicative x-step with the clip of
vmbd_chimney.py:143-147), restated with the FP64 oracle.  Two variants of the
x-step's transposes: 'fixed' (overlap-add extent sw) and 'k1' (the
reference's read of sw+sf-1 columns past each patch, olaGPU.py:305-309,
zeros past the last patch).  The figure is the PSNR against groundtruth.png
after the best affine intensity fit and |shift| <= 3, as in the test.

About 45 s per frame per variant (FP64, scipy):

    python tests/golden/make_chimney_oracle_psnr.py fixed 100
    python tests/golden/make_chimney_oracle_psnr.py k1 100
"""
import json
import os
import sys
import time

import numpy as np
from scipy.optimize import fmin_l_bfgs_b

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import deconv_oracle as DO  # noqa: E402
from oracle import mfbd_oracle as M  # noqa: E402
from oracle import ola_oracle as O  # noqa: E402
from tests.test_gpu_chimney import DATA, aligned_psnr  # noqa: E402
from vmbd_amd.imagetools import imread  # noqa: E402

CHECKPOINTS = (1, 5, 10, 20, 30, 50, 75, 100)


def trajectory(variant, N):
    k1 = variant == 'k1'
    load = lambda i: imread(os.path.join(DATA, 'y_%04d.png' % i)).astype(np.float64)
    gt = imread(os.path.join(DATA, 'groundtruth.png'))
    y0 = sum(load(i).astype(np.float32) for i in range(1, 20)) / np.float32(20)  # vmbd_chimney.py:92-95
    sf = np.array([20, 20])
    csf = (3, 3)
    sx = tuple(np.array(y0.shape) + sf - 1)
    wa = O.Win2WinAux(sx, csf, 0.5)
    x = np.pad(y0.astype(np.float64), ((9, 10), (9, 10)), mode='edge')  # impad(y0, sf - 1)
    sf2 = sf // 2
    out = {}
    t0 = time.time()
    for i in range(1, N + 1):
        y = load(i)
        mask = (y < 1).astype(np.float64)
        X = O.OlaOracle(x, tuple(sf), 'valid', wa)
        sz = (9, 20, 20)
        z0 = X.cnvtp(y).ravel()
        obj = lambda z: DO.cnvinv_objfun(X, z.reshape(sz), y, 0.0, 0.1)
        grad = lambda z: DO.cnvinv_gradfun(X, z.reshape(sz), y, 0.0, 0.1).ravel()
        zhat, _, info = fmin_l_bfgs_b(func=obj, x0=z0, fprime=grad, factr=10., pgtol=10e-15, maxfun=500,
                                      bounds=[(0., np.inf)] * z0.size, m=5)
        f = M.normalize(zhat.reshape(sz))
        F = O.OlaOracle(f, sx, 'valid', wa)
        num = F.cnvtp(mask * y, k1=k1)
        den = F.cnvtp(mask * F.cnv(x), k1=k1)
        x = x * np.maximum(num / (den + 1e-10), 1e-10)
        xmax = x[sf[0]:-sf[0], sf[1]:-sf[1]].max()
        x = np.minimum(x, xmax)
        if i in CHECKPOINTS:
            out[str(i)] = round(float(aligned_psnr(x[sf2[0]:-sf2[0], sf2[1]:-sf2[1]] / xmax, gt)), 3)
            print(f'{variant} frame {i}: psnr {out[str(i)]}  evals {info["funcalls"]}  {time.time() - t0:.0f}s',
                  flush=True)
    return out


def main(variant, N):
    path = os.path.join(HERE, 'chimney_oracle_psnr.json')
    data = json.load(open(path)) if os.path.exists(path) else {}
    data[variant] = trajectory(variant, N)
    with open(path, 'w') as fh:
        json.dump(data, fh, indent=1, sort_keys=True)


if __name__ == '__main__':
    main(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 100)
