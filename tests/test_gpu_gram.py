"""Band overlap matrices on the matrix cores (gh_gram) against FP64
products in torch: C = alpha A B^H (gpaw/utilities/blas.py:39,273,314
semantics), Hermitian lower-triangle mode, conjugated output, ragged band
counts (not multiples of the 128-row blocks) and every dtype.  FP64 /
complex128 within 1e-13 of the exact product; FP32 / complex64 inputs are
summed in FP32 inside k-chunks of up to a few thousand products (a sum of
positive terms, like a diagonal, then carries ~ chunk * 6e-8 relative
error in the worst case), so the bound is 1e-5 of sum |a| |b|."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip('no GPU', allow_module_level=True)

from gpaw_amd import blas  # noqa: E402

DTYPES = [torch.float64, torch.complex128, torch.float32, torch.complex64]


def exact(a, b, alpha, conj):
    a64 = a.to(torch.complex128 if a.is_complex() else torch.float64)
    b64 = b.to(a64.dtype)
    c = alpha * (a64 @ b64.conj().T)
    return c.conj() if conj else c


def bound(a, b):
    """sum_g |a_x(g)| |b_y(g)|: the scale of the rounding error."""
    return (a.abs().to(torch.float64) @ b.abs().to(torch.float64).T)


def tol(dtype):
    return 1e-13 if dtype in (torch.float64, torch.complex128) else 1e-5


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('m,n,k', [(256, 256, 4096), (70, 70, 1728),
                                   (130, 67, 6912), (300, 300, 13824)])
def test_gram_general(dtype, m, n, k):
    g = torch.Generator(device='cuda')
    g.manual_seed(m * 7 + n)
    a = torch.randn((m, k), dtype=dtype, device='cuda', generator=g)
    b = torch.randn((n, k), dtype=dtype, device='cuda', generator=g)
    for conj in (False, True):
        c = blas.gram(a, b, alpha=0.3, conj=conj)
        ref = exact(a, b, 0.3, conj)
        err = (c - ref).abs() / (0.3 * bound(a, b))
        assert float(err.max()) < tol(dtype), (conj, float(err.max()))


@pytest.mark.parametrize('dtype', DTYPES)
@pytest.mark.parametrize('m,k', [(256, 21952), (70, 1728), (300, 4096)])
def test_gram_hermitian(dtype, m, k):
    """S = A A^H (rk) and H = A (HA)^H for a Hermitian H: only the lower
    triangle comes from the data; the result is exactly Hermitian."""
    g = torch.Generator(device='cuda')
    g.manual_seed(m)
    a = torch.randn((m, k), dtype=dtype, device='cuda', generator=g)
    s = blas.gram(a, hermitian=True, alpha=0.5)
    ref = exact(a, a, 0.5, False)
    err = (s - ref).abs() / (0.5 * bound(a, a))
    assert float(err.max()) < tol(dtype)
    assert torch.equal(s, s.conj().T)
    # a second array: B = a M with M Hermitian in the band index
    # would not make A B^H Hermitian; use B = diag(d) A instead
    d = torch.linspace(-1, 2, m, dtype=torch.float64, device='cuda')
    b = (d[:, None].to(a.real.dtype) * a).contiguous()
    h = blas.gram(b, a, hermitian=True, conj=True)
    ref = exact(b, a, 1.0, True)
    lower = torch.tril(torch.ones((m, m), dtype=torch.bool, device='cuda'))
    err = ((h - ref).abs() / bound(b, a))[lower]
    assert float(err.max()) < tol(dtype)
    assert torch.equal(h, h.conj().T)


def test_gram_views_and_errors():
    """Row pitch from a strided band view; the C ABI refuses rows that are
    not 16-byte aligned (the Python layer then takes an FP64 GEMM on the
    GPU, also checked)."""
    from gpaw_amd import _lib
    from gpaw_amd._arrays import stream_ptr
    a = torch.randn((40, 4096), dtype=torch.float64, device='cuda')
    v = a[::2, :2048]
    c = blas.gram(v)
    ref = exact(v, v, 1.0, False)
    assert float(((c - ref).abs() / bound(v, v)).max()) < 1e-13
    w = a[:, 1:2049]              # rows start 8 bytes off a 16-byte boundary
    L = _lib.lib()
    out = torch.empty((40, 40), dtype=torch.float64, device='cuda')
    rc = L.gh_gram(_lib.GH_F64, 40, 40, 2048, w.data_ptr(), 4096,
                   w.data_ptr(), 4096, 1.0, 1, 0, out.data_ptr(), None, 0,
                   stream_ptr())
    assert rc == -4 and b'aligned' in L.gh_last_error()
    c = blas.gram(w, hermitian=True)
    ref = exact(w, w, 1.0, False)
    assert float(((c - ref).abs() / bound(w, w)).max()) < 1e-13
    with pytest.raises(ValueError):
        blas.gram(a[:10], a[:12], hermitian=True)
