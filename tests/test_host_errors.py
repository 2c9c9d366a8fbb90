"""Error behaviour of the drop-in that is decided on the host, before any
device work: exception types and messages of olaGPU.py:104-136 and of the
MFBD/bench argument checks (CPU only)."""
import numpy as np
import pytest

from vmbd_amd.imagetools import win2winaux


@pytest.fixture(scope='module')
def wa():
    return win2winaux((96, 96), (3, 3), 0.5)


def test_olagpu_size_mismatches(wa):
    from vmbd_amd.olaGPU import OlaGPU
    # PSF stack with the wrong number of PSFs (olaGPU.py:109-111)
    with pytest.raises(IOError, match='Size missmatch between winaux and PSF size!'):
        OlaGPU(np.zeros((4, 9, 9), np.float32), (96, 96), 'valid', wa)
    # neither a PSF stack smaller than sx nor an image larger than sf (:120-121)
    with pytest.raises(IOError, match='^Size missmatch$'):
        OlaGPU(np.zeros((9, 120, 120), np.float32), (96, 96), 'valid', wa)
    # image that does not match the windows (:124-125)
    with pytest.raises(IOError, match='Size missmatch between winaux and image size!'):
        OlaGPU(np.zeros((80, 80), np.float32), (9, 9), 'valid', wa)


def test_olagpu_bad_mode(wa):
    from vmbd_amd.olaGPU import OlaGPU
    with pytest.raises(NotImplementedError, match='Not a valid mode!'):
        OlaGPU(np.zeros((9, 9, 9), np.float32), (96, 96), 'wrap', wa)


def test_mfbd_argument_checks():
    from vmbd_amd.mfbd import OnlineMFBD
    y0 = np.zeros((64, 64), np.float32)
    with pytest.raises(ValueError, match='f_solver'):
        OnlineMFBD(y0, (7, 7), f_solver='cg')
    with pytest.raises(ValueError, match='x_solver'):
        OnlineMFBD(y0, (7, 7), x_solver='cg')
    with pytest.raises(ValueError, match='capture'):
        OnlineMFBD(y0, (7, 7), f_solver='lbfgsb', capture=True)
    with pytest.raises(ValueError, match='f_init'):
        OnlineMFBD(y0, (7, 7), f_solver='pbb', f_init='zero')
    with pytest.raises(ValueError, match='prior'):
        OnlineMFBD(y0, (7, 7), f_solver='pbb', prior='tv')


def test_initial_latent_reference_quirk():
    """vmbd_chimney.py:92-95: frames 1..N0-1 summed, divided by N0."""
    from vmbd_amd.mfbd import initial_latent
    frames = [np.full((4, 4), float(i + 1), np.float32) for i in range(30)]
    np.testing.assert_allclose(initial_latent(frames, 20), sum(range(1, 20)) / 20.0)


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, 'bench.py', '--gpus', '2'], capture_output=True, text=True,
                       cwd=__import__('os').path.dirname(__import__('os').path.dirname(__import__('os').path.abspath(__file__))))
    assert r.returncode != 0 and 'one GPU' in r.stderr


@pytest.mark.parametrize('args', [['--streams', '0'], ['--frames', '0']])
def test_bench_refuses_empty_step(args):
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, 'bench.py'] + args, capture_output=True, text=True, cwd=root)
    assert r.returncode == 2 and 'positive counts' in r.stderr


def test_gridf_and_imwrite_roundtrip(tmp_path):
    from vmbd_amd.imagetools import gridF, imread, imwrite
    fs = np.random.default_rng(0).random((9, 5, 5))
    g = gridF(fs, (3, 3))
    assert g.shape == (15, 15)
    np.testing.assert_array_equal(g[5:10, 10:15], fs[5])          # patch k = m*csf1 + n at tile (m, n)
    img = np.linspace(0, 1, 64 * 32).reshape(64, 32)
    imwrite(img, str(tmp_path / 'a.png'))
    back = imread(str(tmp_path / 'a.png'))
    np.testing.assert_array_equal(back * 255, np.floor(img * 255))
