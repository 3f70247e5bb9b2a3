"""BLAS-3 drop-ins on device tensors (gpaw/utilities/blas.py counterparts).

mmm(alpha, a, opa, b, opb, beta, c)     c <- alpha op(a) op(b) + beta c
rk(alpha, a, beta, c, trans='c')        c <- alpha a a^H + beta c
r2k(alpha, a, b, beta, c, trans='c')    c <- alpha b a^H + conj(alpha) a b^H
                                              + beta c (what c/blas.c forms)

Same signatures, argument meaning and shape / dtype checks as the reference
(blas.py:39-87, 273-304, 314-361; c/blas.c:48-170).  Arrays are torch
tensors on the GPU, updated in place; NumPy arrays are staged through the
GPU.  The band-overlap shapes -- a and b with a few hundred rows of a grid's
length, the products rk / r2k / mmm('N', 'C') form for S and H -- run on the
matrix cores (gh_gram: lower triangle for rk, FP64 sums); other shapes are
GEMMs through torch (rocBLAS).  Only the lower triangle of c is defined
after rk and r2k, as in the reference; here c holds the full Hermitian
matrix.
"""
import numpy as np
import torch

from .. import blas as _gblas


def _dev(x):
    if isinstance(x, torch.Tensor):
        if not x.is_cuda:
            raise ValueError('tensor must live on the GPU')
        return x, False
    return torch.from_numpy(np.ascontiguousarray(x)).cuda(), True


def _writeback(orig, t, staged):
    if staged:
        orig[...] = t.cpu().numpy()


def _flat2(a):
    # explicit sizes: torch cannot infer a -1 axis of an empty array
    return a.reshape(a.shape[0], int(np.prod(a.shape[1:])))


def _first_two_flat(a):
    """(n0, n1, rest) view, then the first two axes swapped and the other
    ones combined: (n1, n0 * rest)."""
    rest = int(np.prod(a.shape[2:]))
    a3 = a.reshape(a.shape[0], a.shape[1], rest).permute(1, 0, 2)
    return a3.reshape(a.shape[1], a.shape[0] * rest)


def _tall(a2, b2):
    """Band-overlap shape: long rows (the grid) against few rows."""
    return a2.shape[1] >= 4 * max(a2.shape[0], b2.shape[0])


def _op(x, op):
    return x if op == 'N' else x.T if op == 'T' else x.conj().T


def mmm(alpha, a, opa, b, opb, beta, c):
    """Matrix-matrix multiplication c <- alpha op(a) op(b) + beta c, with op
    'N' (none), 'T' (transpose) or 'C' (conjugate transpose)."""
    assert opa in 'NTC' and opb in 'NTC'
    a1, a2 = a.shape if opa == 'N' else a.shape[::-1]
    b1, b2 = b.shape if opb == 'N' else b.shape[::-1]
    assert a2 == b1
    assert tuple(c.shape) == (a1, b2)
    assert a.dtype == b.dtype == c.dtype
    ta, _ = _dev(a)
    tb, _ = _dev(b)
    tc, staged = _dev(c)
    cplx = tc.is_complex()
    if not cplx:
        assert not isinstance(alpha, complex)
        assert not isinstance(beta, complex)
    if (opa == 'N' and (opb == 'C' or (opb == 'T' and not cplx))
            and _tall(ta, tb)):
        prod = _gblas.gram(ta.contiguous() if ta.stride(1) != 1 else ta,
                           tb.contiguous() if tb.stride(1) != 1 else tb)
        prod = alpha * prod.to(tc.dtype)
    else:
        prod = alpha * (_op(ta, opa) @ _op(tb, opb))
    if beta == 0.0:
        tc.copy_(prod)
    else:
        tc.mul_(beta).add_(prod)
    _writeback(c, tc, staged)


def rk(alpha, a, beta, c, trans='c'):
    """Rank-k update c <- alpha a a^H + beta c (trans='c', the axes after
    the first combined) or c <- alpha a^H a + beta c (trans='n')."""
    assert a.ndim > 1
    assert trans in 'cn'
    ta, _ = _dev(a)
    tc, staged = _dev(c)
    assert ta.is_complex() == tc.is_complex()
    if trans == 'c':
        assert tuple(c.shape) == (a.shape[0], a.shape[0])
        a2 = _flat2(ta)
        if _tall(a2, a2):
            prod = _gblas.gram(a2, hermitian=True).to(tc.dtype)
        else:
            prod = a2 @ a2.conj().T
    else:
        assert tuple(c.shape) == (a.shape[1], a.shape[1])
        a2 = _first_two_flat(ta)
        prod = (a2.conj() @ a2.T)
    prod = alpha * prod
    if beta == 0.0:
        tc.copy_(prod)
    else:
        tc.mul_(beta).add_(prod)
    _writeback(c, tc, staged)


def r2k(alpha, a, b, beta, c, trans='c'):
    """Rank-2k update c <- alpha b a^H + conj(alpha) a b^H + beta c
    (trans='c'; trans='n' contracts over the first axis instead: alpha
    b^H a + conj(alpha) a^H b).  For real alpha, the form GPAW uses, the
    two orders agree."""
    assert trans in 'cn'
    assert a.shape == b.shape and a.dtype == b.dtype
    ta, _ = _dev(a)
    tb, _ = _dev(b)
    tc, staged = _dev(c)
    calpha = np.conj(alpha)
    if trans == 'c':
        assert tuple(c.shape) == (a.shape[0], a.shape[0])
        a2, b2 = _flat2(ta), _flat2(tb)
        if _tall(a2, b2):
            G = _gblas.gram(a2, b2).to(tc.dtype)      # a b^H
        else:
            G = a2 @ b2.conj().T
    else:
        assert tuple(c.shape) == (a.shape[1], a.shape[1])
        a2, b2 = _first_two_flat(ta), _first_two_flat(tb)
        G = a2.conj() @ b2.T                          # a^H b
    # c/blas.c:140-170 hands row-major arrays to zher2k as their Fortran
    # transposes: what it forms is alpha G^H + conj(alpha) G
    prod = alpha * G.conj().T + calpha * G
    if beta == 0.0:
        tc.copy_(prod)
    else:
        tc.mul_(beta).add_(prod)
    _writeback(c, tc, staged)
