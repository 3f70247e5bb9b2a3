"""Band-wise BLAS-1 on device tensors (gpaw/utilities/mblas.py counterparts).

multi_axpy(a_x, x_xG, y_xG): y[x] += a[x] * x[x]       (mblas.py:13-32)
multi_dotc(a_xG, b_xG, scale): Re sum conj(a[x]) b[x] * scale, FP64 result
with FP64 accumulation and a fixed reduction order.
"""
import numpy as np
import torch

from . import _lib
from ._arrays import dtype_code, stream_ptr


def _nb_n(t):
    nb = t.shape[0] if t.dim() > 1 else 1
    return nb, t.numel() // max(nb, 1)


def multi_axpy(a_x, x_xG, y_xG):
    assert x_xG.shape == y_xG.shape and x_xG.dtype == y_xG.dtype
    assert x_xG.is_contiguous() and y_xG.is_contiguous()
    a = torch.as_tensor(a_x, dtype=torch.float64, device='cuda').contiguous()
    nb, n = _nb_n(x_xG)
    assert a.numel() == nb
    _lib.check(_lib.lib().gh_multi_axpy(
        dtype_code(x_xG.dtype), nb, n, a.data_ptr(), x_xG.data_ptr(),
        y_xG.data_ptr(), stream_ptr()))
    return y_xG


def multi_dotc(a_xG, b_xG, scale=1.0, out=None):
    assert a_xG.shape == b_xG.shape and a_xG.dtype == b_xG.dtype
    assert a_xG.is_contiguous() and b_xG.is_contiguous()
    nb, n = _nb_n(a_xG)
    if out is None:
        out = torch.empty(nb, dtype=torch.float64, device='cuda')
    work = torch.empty(nb * _lib.GH_DOT_CHUNKS, dtype=torch.float64,
                       device='cuda')
    _lib.check(_lib.lib().gh_multi_dotc(
        dtype_code(a_xG.dtype), nb, n, a_xG.data_ptr(), b_xG.data_ptr(),
        float(scale), out.data_ptr(), work.data_ptr(), stream_ptr()))
    return out
