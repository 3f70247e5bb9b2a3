"""ctypes binding of libgpawhip.so (C ABI declared in include/gpaw_amd.h).

The library is the only compute path: if it is missing or fails to load, or
no GPU is visible, the calls raise instead of falling back to host code.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'lib', 'libgpawhip.so')

GH_F64, GH_C128, GH_F32, GH_C64 = 0, 1, 2, 3
GH_DOT_CHUNKS = 64

_c = ctypes
P = _c.c_void_p
I64 = _c.c_int64
I32 = _c.c_int32
D = _c.c_double
PI64 = _c.POINTER(_c.c_int64)
PI32 = _c.POINTER(_c.c_int32)
PD = _c.POINTER(_c.c_double)

_SIGNATURES = {
    'gh_last_error': (_c.c_char_p, []),
    'gh_version': (_c.c_char_p, []),
    'gh_stencil_create': (_c.c_int, [_c.POINTER(P), _c.c_int, PD, PI64, PI64,
                                     PI32, _c.c_int]),
    'gh_stencil_destroy': (_c.c_int, [P]),
    'gh_stencil_diagonal': (D, [P]),
    'gh_stencil_apply': (_c.c_int, [P, P, P, I64, PD, P, P, D, P]),
    'gh_stencil_apply_shifted': (_c.c_int, [P, P, P, I64, PD, P, P, P]),
    'gh_stencil_set_tile': (_c.c_int, [P, _c.c_int, _c.c_int, _c.c_int]),
    'gh_stencil_relax': (_c.c_int, [P, _c.c_int, P, P, P, _c.c_int, D, P]),
    'gh_stencil_batch_create': (_c.c_int, [_c.POINTER(P), _c.c_int,
                                           _c.POINTER(P), PI64, PI32, PI32,
                                           _c.c_int]),
    'gh_stencil_batch_destroy': (_c.c_int, [P]),
    'gh_stencil_batch_info': (_c.c_int, [P, PI64, PI32, PI32]),
    'gh_stencil_batch_apply': (_c.c_int, [P, _c.POINTER(P), _c.POINTER(P),
                                          _c.POINTER(P), P]),
    'gh_transformer_create': (_c.c_int, [_c.POINTER(P), PI64, PI64, _c.c_int,
                                         PI64, PI64, PI32, _c.c_int,
                                         _c.c_int]),
    'gh_transformer_destroy': (_c.c_int, [P]),
    'gh_transformer_apply': (_c.c_int, [P, P, P, I64, PD, P]),
    'gh_lfc_create': (_c.c_int, [_c.POINTER(P), I64, _c.c_int, PI64, PI32,
                                 PI32, PI32, PD, _c.c_int, _c.c_int, PD, D,
                                 _c.c_int]),
    'gh_lfc_destroy': (_c.c_int, [P]),
    'gh_lfc_nfunctions': (_c.c_int, [P]),
    'gh_lfc_integrate': (_c.c_int, [P, P, P, I64, _c.c_int, P]),
    'gh_lfc_add': (_c.c_int, [P, P, P, I64, _c.c_int, P]),
    'gh_lfc_apply_dH': (_c.c_int, [P, P, P, P, P, P, I64, P]),
    'gh_stream_copy': (_c.c_int, [P, P, I64, _c.c_int, P]),
    'gh_precond_scratch_bytes': (I64, [P, P, P, I64]),
    'gh_precond_apply': (_c.c_int, [P, P, P, P, P, P, P, D, P, P, I64, PD, P,
                                    P]),
    'gh_multi_axpy': (_c.c_int, [_c.c_int, I64, I64, P, P, P, P]),
    'gh_multi_dotc': (_c.c_int, [_c.c_int, I64, I64, P, P, D, P, P, P]),
    'gh_multi_gram_work': (I64, [_c.c_int, I64]),
    'gh_add_to_density': (_c.c_int, [_c.c_int, I64, I64, P, P, P, P]),
    'gh_multi_gram': (_c.c_int, [_c.c_int, I64, I64, _c.c_int,
                                 _c.POINTER(P), D, P, P, P]),
    'gh_multi_gram_noalloc': (_c.c_int, [_c.c_int, I64, I64, _c.c_int,
                                         _c.POINTER(P), D, P, P]),
    'gh_multi_lincomb': (_c.c_int, [_c.c_int, I64, I64, _c.c_int,
                                    _c.POINTER(P), P, P, P]),
    'gh_gram_work_bytes': (I64, [_c.c_int, I64, I64, I64, _c.c_int]),
    'gh_gram': (_c.c_int, [_c.c_int, I64, I64, I64, P, I64, P, I64, D,
                           _c.c_int, _c.c_int, P, P, I64, P]),
    'gh_unet_conv3_weight_elems': (I64, [_c.c_int, _c.c_int]),
    'gh_unet_pack_conv3': (_c.c_int, [_c.c_int, _c.c_int, P, P]),
    'gh_unet_conv3': (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _c.c_int,
                                 _c.c_int, PI64, P, P, P, P, P, D, P, P,
                                 _c.c_int, P, P]),
    'gh_unet_pool2': (_c.c_int, [_c.c_int, _c.c_int, PI64, P, P, P]),
    'gh_unet_map1': (_c.c_int, [_c.c_int, _c.c_int, I64, P, P, P, P, P]),
    'gh_unet_head': (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, I64, P, P, D,
                                P, P, P]),
    'gh_spline_coefs': (_c.c_int, [_c.c_int, D, PD, PD]),
    'gh_spline_to_grid': (I64, [_c.c_int, _c.c_int, D, PD, PI64, PI64, PD,
                                PD, PI64, PI64, PD, PI32]),
}

_lib = None


class GPAWAMDError(RuntimeError):
    pass


def lib():
    """Load libgpawhip.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GPAWAMDError(
                'libgpawhip.so is not built (expected %s); run '
                'python -m gpaw_amd.build' % LIB_PATH)
        L = _c.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().gh_last_error().decode(errors='replace')
        if rc == -1 or rc == -2:
            raise ValueError('gpaw_amd: %s (code %d)' % (msg, rc))
        raise GPAWAMDError('gpaw_amd: %s (code %d)' % (msg, rc))
    return rc


def i64(a):
    a = np.ascontiguousarray(a, np.int64)
    return a, a.ctypes.data_as(PI64)


def i32(a):
    a = np.ascontiguousarray(a, np.int32)
    return a, a.ctypes.data_as(PI32)


def f64(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(PD)


def phase_arg(phase_cd):
    """Host (3, 2) complex phases -> (keepalive, double*) or (None, None)."""
    if phase_cd is None:
        return None, None
    ph = np.ascontiguousarray(np.asarray(phase_cd, complex).reshape(3, 2))
    return ph, ph.ctypes.data_as(PD)
