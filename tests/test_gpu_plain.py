"""GPU parity of the plain per-operator form (the comparator of the fused
operator) against the FP64 oracle's reference-local transposes."""
import numpy as np
import pytest

from oracle import ola_oracle as O
from tests.helpers import case_inputs, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('sx,csf,sf', [((96, 96), (3, 3), (9, 9)), ((120, 100), (3, 4), (6, 8))])
def test_plain_step_vs_oracle(cuda, sx, csf, sf):
    import torch
    from vmbd_amd.imagetools import win2winaux
    from vmbd_amd.plain import PlainOla
    print(f'case plain step {sx} {csf} {sf}', flush=True)
    x, f, g = case_inputs(41, sx, csf, sf)
    wa = win2winaux(sx, csf, 0.5)
    owa = O.Win2WinAux(sx, csf, 0.5)
    p = PlainOla(sx, sf, wa)
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float32), device=cuda)
    y, xb, fb = p.step(t(x), t(f), t(g))
    x32, f32, g32 = (np.asarray(a, np.float32).astype(np.float64) for a in (x, f, g))
    assert rel_l2(y.cpu().numpy(), O.forward(x32, f32, owa)) < 1e-5
    assert rel_l2(xb.cpu().numpy(), O.adjoint_x(g32, f32, owa, sx, exact=False)) < 1e-5
    assert rel_l2(fb.cpu().numpy(), O.adjoint_f(g32, x32, owa, sf, exact=False)) < 1e-5
