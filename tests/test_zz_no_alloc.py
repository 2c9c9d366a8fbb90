"""The allocation-free PBB step (vmbd_pbb_step_noalloc, static device
scratch) against vmbd_pbb_step with a caller-supplied partials buffer,
called through the C ABI: two calls in a row, the second over fewer
elements, must give bitwise the same iterate and step length."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_pbb_step_noalloc_matches_workspace_form(cuda):
    import torch
    from vmbd_amd import _lib
    L = _lib.lib()
    fn = L.vmbd_pbb_step_noalloc
    rng = np.random.default_rng(91)
    st = _lib.stream_ptr()
    for call, (n, it, unc) in enumerate([(300001, 1, 0), (1003, 2, 1)]):
        print(f'case noalloc call {call}: n={n} iter={it} unconstrained={unc}', flush=True)
        x, ox, g, og = (torch.as_tensor(rng.standard_normal(n).astype(np.float32), device=cuda) for _ in range(4))
        x.clamp_(min=0.0)
        want = torch.empty_like(x)
        got = torch.full_like(x, float('nan'))
        part = torch.empty(3 * L.vmbd_reduce_blocks(n), dtype=torch.float64, device=cuda)
        s_want = torch.empty(1, dtype=torch.float32, device=cuda)
        s_got = torch.full((1,), float('nan'), dtype=torch.float32, device=cuda)
        _lib.call('vmbd_pbb_step', _lib.ptr(want), _lib.ptr(x), _lib.ptr(ox), _lib.ptr(g), _lib.ptr(og), n, it, unc,
                  _lib.ptr(part), _lib.ptr(s_want), st)
        _lib.check(fn(_lib.ptr(got), _lib.ptr(x), _lib.ptr(ox), _lib.ptr(g), _lib.ptr(og), n, it, unc,
                      _lib.ptr(s_got), st), 'vmbd_pbb_step_noalloc')
        torch.cuda.synchronize()
        assert torch.equal(got, want), call
        assert torch.equal(s_got, s_want), call
        assert torch.isfinite(s_got).all()
    # invalid sizes are refused before anything is launched
    assert fn(None, None, None, None, None, 0, 1, 0, None, st) == _lib.VMBD_E_ARG
