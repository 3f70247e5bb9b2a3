"""Grid transformers (drop-in for gpaw.transformers).

Transformer(gdin, gdout, nn, dtype) restricts (gdin twice as fine as gdout)
or interpolates (gdout twice as fine) with 2*nn-point separable filters
(gpaw/transformers.py:18-134).  Ghost widths and skips are computed as in
transformers.py:27-47; the transform runs in libgpawhip
(gh_transformer_*).
"""
import ctypes

import numpy as np

from . import _lib
from ._arrays import Staged, dtype_code, stream_ptr, torch_dtype


class _Transformer:
    def __init__(self, gdin, gdout, nn=1, dtype=float, xp=None):
        self.gdin = gdin
        self.gdout = gdout
        self.nn = nn
        assert 1 <= nn <= 4
        self.dtype = dtype
        self.xp = xp
        pad_cd = np.empty((3, 2), int)
        skip_cd = np.zeros((3, 2), int)
        if (gdin.N_c == 2 * gdout.N_c).all():
            pad_cd[:, 0] = 2 * nn - 1 - 2 * gdout.beg_c + gdin.beg_c
            pad_cd[:, 1] = 2 * nn - 2 + 2 * gdout.end_c - gdin.end_c
            self.interpolate = False
        else:
            assert (gdout.N_c == 2 * gdin.N_c).all()
            pad_cd[:, 0] = nn - 1 - gdout.beg_c // 2 + gdin.beg_c
            pad_cd[:, 1] = nn + gdout.end_c // 2 - gdin.end_c
            skip_cd[:, 0] = gdout.beg_c % 2
            skip_cd[:, 1] = gdout.end_c % 2
            self.interpolate = True
            inpoints = (gdin.n_c[0] + 2 * nn - 1) * (gdin.n_c[1] + 2 * nn - 1)
            outpoints = gdout.n_c[0] * gdout.n_c[1]
            if inpoints > outpoints:
                points = ' x '.join([str(N) for N in gdin.N_c])
                raise ValueError('Cannot construct interpolator.  Grid %s '
                                 'may be too small' % points)
        assert (pad_cd.ravel() >= 0).all()
        self.ngpin = tuple(int(n) for n in gdin.n_c)
        self.ngpout = tuple(int(n) for n in gdout.n_c)
        self.pad_cd = pad_cd
        self.skip_cd = skip_cd
        L = _lib.lib()
        h = ctypes.c_void_p()
        keep = [_lib.i64(gdin.n_c), _lib.i64(gdout.n_c),
                _lib.i64(pad_cd.ravel()), _lib.i64(skip_cd.ravel()),
                _lib.i32(gdin.pbc_c.astype(np.int32))]
        _lib.check(L.gh_transformer_create(
            ctypes.byref(h), keep[0][1], keep[1][1], nn, keep[2][1],
            keep[3][1], keep[4][1], int(self.interpolate),
            dtype_code(dtype)))
        self.handle = h
        self._lib = L

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h is not None and h.value:
            self._lib.gh_transformer_destroy(h)
            self.handle = None

    def apply(self, input, output=None, phases=None):
        if tuple(input.shape[-3:]) != self.ngpin:
            raise ValueError('input grid %s != %s' % (tuple(input.shape[-3:]),
                                                       self.ngpin))
        if output is None:
            output = self.gdout.empty(tuple(input.shape[:-3]), self.dtype)
            if not hasattr(input, 'data_ptr'):
                output = output.cpu().numpy()
        if tuple(output.shape[-3:]) != self.ngpout:
            raise ValueError('output grid %s != %s' % (
                tuple(output.shape[-3:]), self.ngpout))
        if tuple(output.shape[:-3]) != tuple(input.shape[:-3]):
            raise ValueError('input holds %s arrays, output %s' % (
                tuple(input.shape[:-3]), tuple(output.shape[:-3])))
        nb = int(np.prod(input.shape[:-3], dtype=np.int64))
        ph, php = _lib.phase_arg(phases)
        with Staged(input, dtype=self.dtype) as a, \
                Staged(output, writeback=True, dtype=self.dtype) as b:
            _lib.check(self._lib.gh_transformer_apply(
                self.handle, a.data_ptr(), b.data_ptr(), nb, php,
                stream_ptr()))
        return output


def Transformer(gdin, gdout, nn=1, dtype=float, xp=None):
    if nn != 9:
        return _Transformer(gdin, gdout, nn, dtype, xp)

    class T:
        nn = 1

        def apply(self, input, output, phases=None):
            output[:] = input
    return T()
