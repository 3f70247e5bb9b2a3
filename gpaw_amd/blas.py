"""Band-wise BLAS-1 on device tensors (gpaw/utilities/mblas.py counterparts).

multi_axpy(a_x, x_xG, y_xG): y[x] += a[x] * x[x]       (mblas.py:13-32)
multi_dotc(a_xG, b_xG, scale): Re sum conj(a[x]) b[x] * scale, FP64 result
with FP64 accumulation and a fixed reduction order.
multi_gram(vecs, scale): G[x, i, j] = scale * sum conj(v_i[x]) v_j[x] for a
list of k <= 6 arrays of equal shape (one pass over the data).
multi_lincomb(vecs, coef, out): out[x] = sum_i coef[x, i] vecs[i][x] for
k <= 8 arrays (one pass; `out` may be one of the inputs).
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


def multi_gram(vecs, scale=1.0):
    """Per-band Gram matrices (nb, k, k): float64 for real inputs, complex128
    for complex ones."""
    k = len(vecs)
    v0 = vecs[0]
    for v in vecs:
        assert v.shape == v0.shape and v.dtype == v0.dtype
        assert v.is_contiguous() and v.is_cuda
    nb, n = _nb_n(v0)
    L = _lib.lib()
    cplx = v0.is_complex()
    out = torch.empty((nb, k, k, 2) if cplx else (nb, k, k),
                      dtype=torch.float64, device='cuda')
    nwork = L.gh_multi_gram_work(k, nb)
    if nwork < 0:
        raise ValueError('multi_gram: k = %d out of range' % k)
    work = torch.empty(max(int(nwork), 1), dtype=torch.float64,
                       device='cuda')
    ptrs = (_lib.P * k)(*[v.data_ptr() for v in vecs])
    _lib.check(L.gh_multi_gram(
        dtype_code(v0.dtype), nb, n, k, ptrs, float(scale), out.data_ptr(),
        work.data_ptr(), stream_ptr()))
    return torch.view_as_complex(out) if cplx else out


def multi_lincomb(vecs, coef, out=None):
    """out[x] = sum_i coef[x, i] vecs[i][x], terms accumulated in the order
    given (rmmdiis.py:226-237; with coef = (1, lam) the out-of-place
    per-band axpy).  coef: (nbands, k) real or complex; complex
    coefficients need complex data."""
    k = len(vecs)
    v0 = vecs[0]
    for v in vecs:
        assert v.shape == v0.shape and v.dtype == v0.dtype
        assert v.is_contiguous() and v.is_cuda
    nb, n = _nb_n(v0)
    if out is None:
        out = torch.empty_like(v0)
    assert out.shape == v0.shape and out.dtype == v0.dtype
    assert out.is_contiguous()
    coef = torch.as_tensor(coef, device='cuda')
    assert coef.shape == (nb, k), coef.shape
    if v0.is_complex():
        c = torch.view_as_real(coef.to(torch.complex128).contiguous())
    else:
        assert not coef.is_complex(), 'complex coefficients for real data'
        c = coef.to(torch.float64)
    c = c.contiguous()
    ptrs = (_lib.P * k)(*[v.data_ptr() for v in vecs])
    _lib.check(_lib.lib().gh_multi_lincomb(
        dtype_code(v0.dtype), nb, n, k, ptrs, c.data_ptr(), out.data_ptr(),
        stream_ptr()))
    return out


def gram(a_xG, b_yG=None, alpha=1.0, hermitian=False, conj=False, out=None):
    """C[x, y] = alpha * sum_G a[x, G] conj(b[y, G]) on the matrix cores
    (gh_gram): float64 result for real data, complex128 for complex data,
    whatever the input precision (FP32 / complex64 products are summed in
    FP32 within k-chunks and the chunks in FP64).  b_yG=None means a_xG.
    hermitian: the product is known to be Hermitian (m == n); only the lower
    triangle is formed from the data and the upper one is its conjugate.
    conj: return the conjugate, sum conj(a) b."""
    if b_yG is None:
        b_yG = a_xG
    assert a_xG.dtype == b_yG.dtype and a_xG.is_cuda and b_yG.is_cuda
    def rows(t):
        if t.dim() == 1:
            return t[None]
        return t.reshape(t.shape[0], int(np.prod(t.shape[1:])))
    a, b = rows(a_xG), rows(b_yG)
    assert a.shape[1] == b.shape[1], (a.shape, b.shape)
    assert a.stride(1) == 1 and b.stride(1) == 1
    m, n, k = a.shape[0], b.shape[0], a.shape[1]
    if hermitian and m != n:
        raise ValueError('hermitian product needs equal row counts')
    cplx = a.is_complex()
    shape = (m, n)
    if out is None:
        out = torch.empty(shape, dtype=torch.complex128 if cplx else
                          torch.float64, device='cuda')
    assert tuple(out.shape) == shape and out.is_contiguous()
    assert out.dtype == (torch.complex128 if cplx else torch.float64)
    es = a.element_size()
    if ((k * es) % 16 or (a.stride(0) * es) % 16 or (b.stride(0) * es) % 16
            or a.data_ptr() % 16 or b.data_ptr() % 16):
        # rows the 16-byte matrix-core loader cannot take (odd grid sizes,
        # misaligned views): FP64 GEMM on the GPU
        mdt = out.dtype
        c = alpha * (a.to(mdt) @ b.to(mdt).conj().T)
        if hermitian:
            low = torch.tril(c, -1)
            c = low + low.conj().T + torch.diag(torch.diagonal(c).real).to(mdt)
        out.copy_(c.conj() if conj else c)
        return out
    L = _lib.lib()
    code = dtype_code(a.dtype)
    nwork = int(L.gh_gram_work_bytes(code, m, n, k, int(hermitian)))
    work = torch.empty(max(nwork, 1), dtype=torch.uint8, device='cuda')
    _lib.check(L.gh_gram(
        code, m, n, k, a.data_ptr(), a.stride(0), b.data_ptr(), b.stride(0),
        float(alpha), int(hermitian), int(conj), out.data_ptr(),
        work.data_ptr(), nwork, stream_ptr()))
    return out
