"""Host-side logic of the product (CPU only): window tables handed to the
kernels, geometry helpers, and the C ABI exports (no compute calls)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import ola_oracle as O
from tests.helpers import GEOMS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize('geom', GEOMS + [((1024, 1024), (10, 10), (41, 41), 0.5),
                                          ((768, 1024), (6, 6), (25, 25), 0.5)], ids=str)
def test_product_windows_match_oracle(geom):
    from vmbd_amd.imagetools import win2winaux
    sx, csf, sf, ov = geom
    wa = win2winaux(sx, csf, ov)
    owa = O.Win2WinAux(sx, csf, ov)
    assert tuple(wa.sw) == tuple(owa.sw) and tuple(wa.nhop) == tuple(owa.nhop)
    np.testing.assert_allclose(wa.ws, owa.ws, rtol=2e-7, atol=1e-20)
    # the class table reproduces the full stack exactly
    np.testing.assert_array_equal(wa.win_table[wa.win_idx], wa.ws)
    assert wa.win_table.shape[0] <= 9 or ov != 0.5


def test_fft_tile():
    from vmbd_amd.imagetools import fft_tile
    assert fft_tile((188, 188), (41, 41)) == (256, 256)
    assert fft_tile((220, 293), (25, 25)) == (256, 512)
    assert fft_tile((3, 3), (3, 3)) == (16, 16)


def _header_functions():
    src = ''.join(open(os.path.join(ROOT, 'include', f)).read()
                  for f in sorted(os.listdir(os.path.join(ROOT, 'include'))) if f.endswith('.h'))
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(vmbd_[a-z0-9_]+)\s*\(', src)))


def test_library_exports_every_declared_symbol():
    lib = os.path.join(ROOT, 'vmbd_amd', 'libvmbd_hip.so')
    if not os.path.exists(lib):
        pytest.skip('library not built (run python -m vmbd_amd.build)')
    names = _header_functions()
    assert len(names) >= 15
    out = subprocess.run(['nm', '-D', '--defined-only', lib], capture_output=True, text=True).stdout
    exported = set(ln.split()[-1] for ln in out.splitlines() if ln.strip())
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # the library loads (binding to the HIP runtime torch loads) and the
    # Python signature table covers the header
    import vmbd_amd._lib as L
    h = L.lib()
    for n in names:
        assert hasattr(h, n)
    assert set(names) <= set(L.SIGNATURES)
    assert h.vmbd_version() >= 100


def test_plan_create_rejects_bad_geometry_without_gpu_work():
    """Argument validation happens on the host before any allocation."""
    import vmbd_amd._lib as L
    lib = os.path.join(ROOT, 'vmbd_amd', 'libvmbd_hip.so')
    if not os.path.exists(lib):
        pytest.skip('library not built')
    h = L.lib()
    g = L.VmbdGeom(64, 64, 5, 5, 2, 2, 32, 32, 16, 16, 100, 64, 0)   # tile not a power of two
    tab = np.ones((1, 32, 32), np.float32)
    idx = np.zeros(4, np.int32)
    p = h.vmbd_plan_create(ctypes.byref(g), tab.ctypes.data_as(ctypes.c_void_p), 1, idx.ctypes.data_as(ctypes.c_void_p))
    assert not p and b'power' in h.vmbd_last_error()
    g = L.VmbdGeom(64, 64, 5, 5, 2, 2, 32, 32, 16, 16, 32, 64, 0)    # tile < sw+sf-1
    p = h.vmbd_plan_create(ctypes.byref(g), tab.ctypes.data_as(ctypes.c_void_p), 1, idx.ctypes.data_as(ctypes.c_void_p))
    assert not p and b'smaller' in h.vmbd_last_error()
    g = L.VmbdGeom(64, 64, 5, 5, 2, 2, 32, 32, 16, 16, 64, 64, 0)    # grid covers 48 < 64
    p = h.vmbd_plan_create(ctypes.byref(g), tab.ctypes.data_as(ctypes.c_void_p), 1, idx.ctypes.data_as(ctypes.c_void_p))
    assert not p and b'covers' in h.vmbd_last_error()
    idx[2] = 5
    g = L.VmbdGeom(64, 64, 5, 5, 2, 2, 48, 48, 24, 24, 64, 64, 0)
    p = h.vmbd_plan_create(ctypes.byref(g), tab.ctypes.data_as(ctypes.c_void_p), 1, idx.ctypes.data_as(ctypes.c_void_p))
    assert not p and b'win_idx' in h.vmbd_last_error()


def test_mixed_batch_stream_lanes():
    """Groups go to streams largest first, each to the stream with the fewest
    pixels so far; every group lands on exactly one stream."""
    import torch
    from vmbd_amd.batch import MixedBatch

    def lanes(streams, groups):
        mb = MixedBatch.__new__(MixedBatch)
        mb.streams = streams
        mb.groups = [dict(idx=list(range(n)), x=torch.zeros((n,) + sx)) for n, sx in groups]
        work = lambda gi: len(mb.groups[gi]['idx']) * mb.groups[gi]['x'][0].numel()
        return mb._lanes(list(range(len(groups))), work)

    # pixels per group: 4*16, 2*64, 1*16, 3*16, 1*64
    groups = [(4, (4, 4)), (2, (8, 8)), (1, (4, 4)), (3, (4, 4)), (1, (8, 8))]
    assert lanes(1, groups) == [[1, 0, 4, 3, 2]]
    two = lanes(2, groups)
    assert two == [[1, 3], [0, 4, 2]]
    assert sorted(g for lane in two for g in lane) == list(range(5))
    # more streams than groups: one group per stream
    assert lanes(8, groups[:3]) == [[1], [0], [2]]
    with pytest.raises(ValueError):
        MixedBatch([], streams=0)
