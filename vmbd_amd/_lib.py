"""ctypes binding of libvmbd_hip.so (declarations in include/vmbd.h).

torch is imported first so that the library binds to the HIP runtime torch
already loaded (both carry the soname libamdhip64.so.7).  A missing library
is an error: there is no fallback path.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede the CDLL load, see module docstring)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libvmbd_hip.so')

VMBD_OK = 0
MODES = {'valid': 0, 'same': 1, 'full': 2}
CACHE_PSF = 0
CACHE_IMAGE = 1


REUSE_ROWS = 1


class VmbdTuning(ctypes.Structure):
    _fields_ = [('ola_threads', ctypes.c_int32), ('col_threads', ctypes.c_int32), ('image_rows', ctypes.c_int32)]


class VmbdGeom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        'H', 'W', 'sfh', 'sfw', 'csf0', 'csf1', 'sw0', 'sw1',
        'nhop0', 'nhop1', 'Th', 'Tw', 'mode', 'max_frames')]


_lib = None

_P = ctypes.c_void_p
_F = ctypes.c_void_p  # device float*
_I = ctypes.c_int32

# name -> (restype, argtypes); every name is exported by the library
SIGNATURES = {
    'vmbd_version': (ctypes.c_int, []),
    'vmbd_last_error': (ctypes.c_char_p, []),
    'vmbd_last_status': (ctypes.c_int, []),
    'vmbd_plan_create': (_P, [ctypes.POINTER(VmbdGeom), ctypes.c_void_p, _I, ctypes.c_void_p]),
    'vmbd_plan_destroy': (None, [_P]),
    'vmbd_plan_set_tuning': (ctypes.c_int, [_P, ctypes.POINTER(VmbdTuning)]),
    'vmbd_plan_get_tuning': (ctypes.c_int, [_P, ctypes.POINTER(VmbdTuning)]),
    'vmbd_plan_set_window_factors': (ctypes.c_int, [_P, _P, _P, ctypes.c_float]),
    'vmbd_plan_image_rows_shared': (ctypes.c_int, [_P, _I]),
    'vmbd_plan_workspace_bytes': (ctypes.c_size_t, [_P]),
    'vmbd_plan_out_shape': (ctypes.c_int, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    'vmbd_plan_spectrum_elems': (ctypes.c_size_t, [_P]),
    'vmbd_ola_forward': (ctypes.c_int, [_P, _I, _F, _F, _F, _P]),
    'vmbd_ola_adjoint_x': (ctypes.c_int, [_P, _I, _F, _F, _F, _I, _P]),
    'vmbd_ola_adjoint_f': (ctypes.c_int, [_P, _I, _F, _F, _F, _I, _P]),
    'vmbd_ola_adjoint_xf': (ctypes.c_int, [_P, _I, _F, _F, _F, _F, _F, _I, _I, _P]),
    'vmbd_spectra': (ctypes.c_int, [_P, _I, _I, _F, _P, _P]),
    'vmbd_cache_psf_spectra': (ctypes.c_int, [_P, _I, _F, _P, _P]),
    'vmbd_cache_image_spectra': (ctypes.c_int, [_P, _I, _F, _P, _P]),
    'vmbd_windows': (ctypes.c_int, [_P, _F, _P]),
    'vmbd_ola_forward_cached': (ctypes.c_int, [_P, _I, _I, _P, _F, _F, _P]),
    'vmbd_ola_adjoint_cached': (ctypes.c_int, [_P, _I, _I, _P, _F, _F, _I, _P]),
    'vmbd_ola_forward_cached_ex': (ctypes.c_int, [_P, _I, _I, _P, _F, _F, _F, _F, _P]),
    'vmbd_ola_adjoint_cached_ex': (ctypes.c_int, [_P, _I, _I, _P, _F, _F, ctypes.c_float, _F, _I, _P]),
    'vmbd_batch_create': (_P, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _I]),
    'vmbd_batch_destroy': (None, [_P]),
    'vmbd_batch_launches': (ctypes.c_int, [_P]),
    'vmbd_batch_step': (ctypes.c_int, [_P, _P]),
    'vmbd_deconv_workspace_bytes': (ctypes.c_size_t, [_P]),
    'vmbd_deconv_direct': (ctypes.c_int, [_P, _I, _P, _F, ctypes.c_float, ctypes.c_float, _F, _P]),
    'vmbd_deconv_gdirect': (ctypes.c_int, [_P, _F, _F, _F, _F, ctypes.c_float, ctypes.c_float, _F, _P, _P]),
    'vmbd_deconv_sparse_setup': (ctypes.c_int, [_P, _P, _F, _F, ctypes.c_float, ctypes.c_float, _P, _F, _P, _P]),
    'vmbd_deconv_sparse_step': (ctypes.c_int, [_P, _P, _P, _F, _I, _I, _F, _I, _I, ctypes.c_float, ctypes.c_float,
                                               ctypes.c_float, _F, _F, _P, _P]),
    # plain (reference-decomposition) kernels, include/vmbd_plain.h
    'vmbd_pad': (ctypes.c_int, [_F, _I, _I, _F, _I, _I, _I, _I, _I, _P]),
    'vmbd_chop_pad': (ctypes.c_int, [_F, _I, _I, _F, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _F, _F, _I, _P]),
    'vmbd_pad_stack': (ctypes.c_int, [_F, _I, _I, _I, _F, _I, _I, _I, _I, _I, _P]),
    'vmbd_crop_stack': (ctypes.c_int, [_F, _I, _I, _I, _F, _I, _I, _I, _I, _I, _P]),
    'vmbd_crop': (ctypes.c_int, [_F, _I, _I, _F, _I, _I, _I, _I, _I, _P]),
    'vmbd_ola': (ctypes.c_int, [_F, _F, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _P]),
    'vmbd_fft2': (ctypes.c_int, [_F, _I, _I, _I, _I, ctypes.c_float, _F, _P]),
    'vmbd_twiddles': (ctypes.c_int, [_F, _P]),
    'vmbd_cmul': (ctypes.c_int, [_F, _F, _F, ctypes.c_int64, ctypes.c_int64, _I, ctypes.c_float, _P]),
    'vmbd_fourier_div': (ctypes.c_int, [_F, _F, _F, _F, ctypes.c_int64, ctypes.c_int64, ctypes.c_float,
                                        ctypes.c_float, _P]),
    'vmbd_fourier_div_g': (ctypes.c_int, [_F, _F, _F, _F, _F, _F, ctypes.c_int64, ctypes.c_int64, ctypes.c_float,
                                          ctypes.c_float, _P]),
    'vmbd_fourier_div_s': (ctypes.c_int, [_F, _F, _F, _F, _F, _F, _F, _F, ctypes.c_int64, ctypes.c_int64,
                                          ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]),
    # solver kernels, include/vmbd_solver.h
    'vmbd_reduce_blocks': (ctypes.c_int, [ctypes.c_int64]),
    'vmbd_clip': (ctypes.c_int, [_F, ctypes.c_int64, ctypes.c_float, ctypes.c_float, _I, _P]),
    'vmbd_psf_normalize': (ctypes.c_int, [_F, _I, _I, ctypes.c_float, _P]),
    'vmbd_dot': (ctypes.c_int, [_F, _F, ctypes.c_int64, _F, _F, _P]),
    'vmbd_pbb_stats': (ctypes.c_int, [_F, _F, _F, _F, ctypes.c_int64, _I, _F, _F, _P]),
    'vmbd_pbb_update': (ctypes.c_int, [_F, _F, _F, ctypes.c_int64, _F, _I, _I, _F, _P]),
    'vmbd_axpby': (ctypes.c_int, [_F, _F, ctypes.c_float, _F, ctypes.c_float, ctypes.c_int64, _P]),
    'vmbd_mul': (ctypes.c_int, [_F, _F, _F, ctypes.c_int64, _P]),
    'vmbd_ew': (ctypes.c_int, [_F, _F, _F, ctypes.c_float, ctypes.c_int64, _I, _P]),
    'vmbd_pbb_step': (ctypes.c_int, [_F, _F, _F, _F, _F, ctypes.c_int64, _I, _I, _P, _F, _P]),
    'vmbd_pbb_select': (ctypes.c_int, [_F, _F, _F, _F, ctypes.c_int64, _P, ctypes.c_double, _P]),
    'vmbd_pbb_step_noalloc': (ctypes.c_int, [_F, _F, _F, _F, _F, ctypes.c_int64, _I, _I, _F, _P]),
    'vmbd_tv': (ctypes.c_int, [_F, _I, _I, _I, ctypes.c_float, _P, _P, _P]),
    'vmbd_tv_grad': (ctypes.c_int, [_F, _F, _I, _I, _I, ctypes.c_float, _P, ctypes.c_float, _I, _P]),
    'vmbd_rl_update': (ctypes.c_int, [_F, _F, _F, _I, _I, ctypes.c_float, _I, _I, _F, _F, _P]),
    # filters, include/vmbd_filters.h
    'vmbd_shock_step': (ctypes.c_int, [_F, _F, _I, _I, ctypes.c_float, ctypes.c_float, _P]),
    'vmbd_bilateral': (ctypes.c_int, [_F, _F, _I, _I, _I, ctypes.c_float, ctypes.c_float, _P]),
    'vmbd_gradient': (ctypes.c_int, [_F, _F, _F, _I, _I, _I, _P]),
    'vmbd_impad': (ctypes.c_int, [_F, _F, _I, _I, _I, _I, _P]),
    'vmbd_laplace': (ctypes.c_int, [_F, _F, _I, _I, _I, _I, _I, _P]),
    'vmbd_laplace3d': (ctypes.c_int, [_F, _F, _I, _I, _I, _P]),
    'vmbd_shrink_alpha23': (ctypes.c_int, [_F, ctypes.c_int64, ctypes.c_float, _P]),
    'vmbd_shrink_alpha': (ctypes.c_int, [_F, ctypes.c_int64, ctypes.c_float, ctypes.c_float, _P]),
    'vmbd_edgetaper': (ctypes.c_int, [_F, _I, _I, _F, _I, _F, _I, _P]),
    'vmbd_resize': (ctypes.c_int, [_F, _I, _I, _F, _I, _I, _I, _P]),
    'vmbd_prof_enable': (ctypes.c_int, [_I]),
    'vmbd_prof_reset': (ctypes.c_int, []),
    'vmbd_prof_report': (ctypes.c_int, [ctypes.c_char_p, _I]),
}


def lib():
    """The loaded library (loaded on first use)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f'{LIB_PATH} is missing: build it with `python -m vmbd_amd.build` '
                '(the HIP extension is required; there is no CPU fallback)')
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


VMBD_E_ARG = -1
VMBD_E_HIP = -2
VMBD_E_NOMEM = -3
VMBD_E_UNSUP = -4


class VmbdError(RuntimeError):
    """A failed library call; ``code`` is its VMBD_E* status."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def check(rc, what=''):
    if rc != VMBD_OK:
        msg = lib().vmbd_last_error().decode(errors='replace')
        raise VmbdError(f'{what}: vmbd error {rc}: {msg}', rc)


def call(name, *args):
    """Call an entry point and raise VmbdError on a non-zero status."""
    check(getattr(lib(), name)(*args), name)


def last_error():
    return lib().vmbd_last_error().decode(errors='replace')


def wrote(*tensors):
    """Record that the library has written these tensors through raw
    pointers: torch's version counter of each is bumped, as a torch in-place
    op would, so that OlaPlan's row-transform reuse (keyed on the version)
    never trusts a transform of the old contents.  None entries and
    inference tensors (which carry no counter) are skipped."""
    for t in tensors:
        if isinstance(t, torch.Tensor) and not t.is_inference():
            torch.autograd.graph.increment_version(t)


def ptr(t):
    """Device pointer of a contiguous tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def prof_enable(on=True):
    lib().vmbd_prof_enable(1 if on else 0)


def prof_reset():
    lib().vmbd_prof_reset()


def prof_report():
    """{kernel name: (count, total_us, avg_us)} of the recorded launches."""
    L = lib()
    n = L.vmbd_prof_report(None, 0)
    buf = ctypes.create_string_buffer(max(1, n))
    L.vmbd_prof_report(buf, n)
    out = {}
    for ln in buf.value.decode().splitlines():
        name, cnt, tot, avg = ln.split()
        out[name] = (int(cnt), float(tot), float(avg))
    return out
